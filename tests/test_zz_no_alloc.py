"""svh_predict_noalloc (device pointers in and out, working memory from the
library's fixed device array, no allocation) against svh_predict on the same
inputs: two calls in a row, the second smaller, then a call too large for the
fixed array, which must be refused.  GPU only.

The test's own device buffers come from the HIP runtime libsvh.so links
(hipMalloc / hipMemcpy through ctypes), so only one HIP runtime is in the
process."""

import ctypes
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_H2D, _D2H = 1, 2


def _hip():
    from pyscaledvecchia_amd import _lib
    _lib._load()                                  # libsvh.so pulls in its runtime first
    H = ctypes.CDLL("libamdhip64.so.7")
    H.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    H.hipFree.argtypes = [ctypes.c_void_p]
    H.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    H.hipDeviceSynchronize.argtypes = []
    return H


class _Dev:
    def __init__(self, H, nbytes):
        self.H = H
        self.p = ctypes.c_void_p()
        assert H.hipMalloc(ctypes.byref(self.p), max(int(nbytes), 8)) == 0
        self.nbytes = nbytes

    @classmethod
    def from_array(cls, H, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        d = cls(H, a.nbytes)
        assert H.hipMemcpy(d.p, a.ctypes.data, a.nbytes, _H2D) == 0
        return d

    def to_array(self, n):
        out = np.empty(n)
        assert self.H.hipMemcpy(out.ctypes.data, self.p, out.nbytes, _D2H) == 0
        return out

    def free(self):
        self.H.hipFree(self.p)


def _case(seed, ntr, nq, d):
    rng = np.random.default_rng(seed)
    X = rng.random((ntr, d))
    resid = np.sin(4 * X[:, 0]) + X[:, 1] * X[:, -1]
    Xq = rng.random((nq, d))
    theta = np.concatenate([[0.2], np.log(np.linspace(0.3, 0.9, d)), [math.log(1e-3)]])
    return X, resid, Xq, theta


def test_predict_noalloc_matches_predict(gpu):
    H = _hip()
    for seed, ntr, nq, d, m, nf in ((1, 3000, 400, 4, 140, True), (2, 900, 37, 3, 30, False)):
        print(f"case ntr={ntr} nq={nq} d={d} m={m}", flush=True)
        X, resid, Xq, theta = _case(seed, ntr, nq, d)
        ref_mu, ref_va = gpu.predict(X, resid, Xq, m, theta, 2.5, noise_free=nf)
        bufs = [_Dev.from_array(H, a) for a in (X, resid, Xq, np.full(nq, np.nan),
                                                np.full(nq, np.nan))]
        dX, dr, dq, dmu, dva = bufs
        gpu.predict_noalloc(dX.p.value, dr.p.value, ntr, dq.p.value, nq, d, m, theta, 2.5,
                            dmu.p.value, dva.p.value, noise_free=nf)
        mu, va = dmu.to_array(nq), dva.to_array(nq)
        for b in bufs:
            b.free()
        # same kernels and launch schedule as svh_predict: identical results
        np.testing.assert_array_equal(mu, ref_mu)
        np.testing.assert_array_equal(va, ref_va)


def test_predict_noalloc_refuses_oversize(gpu):
    H = _hip()
    ntr, nq, d, m = 20000, 60000, 4, 140     # 8 nq m alone is 67 MB > 64 MiB
    t = _Dev(H, 64)
    with pytest.raises(ValueError, match="fixed device"):
        gpu.predict_noalloc(t.p.value, t.p.value, ntr, t.p.value, nq, d, m, np.zeros(d + 2), 2.5,
                            t.p.value, t.p.value)
    t.free()
