"""ctypes binding of libsvh.so (the gfx950 HIP library, include/svh.h).

There is no CPU fallback: if the library is missing or no HIP device is
visible, every call raises.
"""

from __future__ import annotations

import ctypes
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsvh.so")

SVH_OK, SVH_ERR_CHOLESKY, SVH_ERR_ARG, SVH_ERR_HIP = 0, 1, 2, 3
PRECISION = {"fp64": 0, "mixed": 1, "mixed_f32cov": 2}

_lib = None
_lock = threading.Lock()
_ctx = None

_dp = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)
_i32p = ctypes.POINTER(ctypes.c_int32)
_vp = ctypes.c_void_p


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built; run `make -C pyscaledvecchia_amd/csrc` "
            "or __graft_entry__.build()")
    L = ctypes.CDLL(LIB_PATH)
    L.svh_device_count.restype = ctypes.c_int
    L.svh_version.restype = ctypes.c_char_p
    L.svh_create.argtypes = [ctypes.c_int, ctypes.POINTER(_vp)]
    L.svh_destroy.argtypes = [_vp]
    L.svh_destroy.restype = None
    L.svh_last_error.argtypes = [_vp]
    L.svh_last_error.restype = ctypes.c_char_p
    L.svh_maximin.argtypes = [_vp, _dp, ctypes.c_int64, ctypes.c_int32, ctypes.c_int64, _i64p]
    L.svh_ordered_knn.argtypes = [_vp, _dp, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32,
                                  ctypes.c_int64, _i64p]
    L.svh_ordered_knn_rows.argtypes = [_vp, _dp, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32,
                                       ctypes.c_int64, _i64p]
    L.svh_knn_query.argtypes = [_vp, _dp, ctypes.c_int64, _dp, ctypes.c_int64, ctypes.c_int32,
                                ctypes.c_int32, _i64p]
    L.svh_vecchia_setup.argtypes = [_vp, _dp, _dp, _dp, ctypes.c_int64, ctypes.c_int32,
                                    ctypes.c_int32, ctypes.c_int64, _i64p, _i64p, _i32p,
                                    ctypes.c_int32, _i32p, _i32p, ctypes.c_int32, ctypes.c_int32]
    L.svh_vecchia_rebuild.argtypes = [_vp, _dp, _dp, _dp, ctypes.c_int64, ctypes.c_int32,
                                      ctypes.c_int32, _dp, ctypes.c_int64, ctypes.c_int32,
                                      ctypes.c_int32, _i64p]
    L.svh_vecchia_eval.argtypes = [_vp, _dp, ctypes.c_int32, ctypes.c_double, ctypes.c_int32,
                                   ctypes.c_int32, ctypes.c_double, ctypes.c_double, _dp, _dp,
                                   _dp, _dp, _i32p]
    L.svh_predict.argtypes = [_vp, _dp, _dp, ctypes.c_int64, _dp, ctypes.c_int64, ctypes.c_int32,
                              ctypes.c_int32, _dp, ctypes.c_double, ctypes.c_int32,
                              ctypes.c_int32, _dp, _dp, _i64p]
    L.svh_predict_noalloc.argtypes = [_vp, _vp, _vp, ctypes.c_int64, _vp, ctypes.c_int64,
                                      ctypes.c_int32, ctypes.c_int32, _dp, ctypes.c_double,
                                      ctypes.c_int32, ctypes.c_int32, _vp, _vp]
    L.svh_matern_general.argtypes = [ctypes.c_double, _dp, ctypes.c_int64, _dp, _dp]
    L.svh_matern_general.restype = ctypes.c_int
    L.svh_block_columns.argtypes = [_vp, _dp, ctypes.c_int64, ctypes.c_int32, _i64p, _dp,
                                    ctypes.c_int64, ctypes.c_int32, _dp, ctypes.c_double,
                                    ctypes.c_int32, _dp]
    L.svh_timing_enable.argtypes = [_vp, ctypes.c_int32]
    L.svh_timing_read.argtypes = [_vp, ctypes.c_char_p, _dp, _i64p]
    L.svh_timing_reset.argtypes = [_vp]
    L.svh_ordering_batch.argtypes = [_vp, ctypes.c_int32, _i64p, _i32p, _i64p, _dp,
                                     ctypes.c_int32, _i64p, _i64p]
    L.svh_batch_setup.argtypes = [_vp, ctypes.c_int32, _i32p, _i32p, _i32p, _i64p, _dp, _dp, _dp,
                                  _i64p, _i32p, _i32p, _i32p, ctypes.c_int32, _dp, _i32p, _dp,
                                  ctypes.c_double]
    L.svh_batch_update.argtypes = [_vp, ctypes.c_int32, _dp, _dp, _dp, _i64p, _i32p]
    L.svh_batch_data.argtypes = [_vp, _dp, _dp, _dp]
    L.svh_joint_solve.argtypes = [_vp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int32, _i64p, _dp,
                                  _dp, ctypes.c_int32, _dp, _dp, _dp, ctypes.c_int32, _dp]
    L.svh_batch_rebuild.argtypes = [_vp, ctypes.c_int32, _i32p, _dp, _i64p, _i32p]
    L.svh_batch_fit_start.argtypes = [_vp, _dp, ctypes.c_int32, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_int32]
    L.svh_batch_fit_step.argtypes = [_vp, ctypes.c_int32, _i32p]
    L.svh_batch_get.argtypes = [_vp, _dp, _dp, _dp, _i32p, _i32p, _i32p]
    L.svh_batch_set_reorder.argtypes = [_vp, ctypes.c_int32, ctypes.c_int32]
    L.svh_batch_eval.argtypes = [_vp, _dp, _dp, _dp, _dp, _dp, _i32p]
    _fp = ctypes.POINTER(ctypes.c_float)
    L.svh_mlp_init.argtypes = [_vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                               ctypes.c_int32, ctypes.c_int32]
    L.svh_mlp_num_params.argtypes = [_vp, _i64p]
    L.svh_mlp_set_params.argtypes = [_vp, _fp]
    L.svh_mlp_get_params.argtypes = [_vp, _fp]
    L.svh_mlp_forward.argtypes = [_vp, _fp, ctypes.c_int32, _fp]
    L.svh_mlp_loss_grad.argtypes = [_vp, _fp, _fp, _fp, ctypes.c_int32, _fp, _fp]
    L.svh_mlp_load_data.argtypes = [_vp, _fp, _fp, _fp, ctypes.c_int32]
    L.svh_mlp_train_epoch.argtypes = [_vp, _i32p, ctypes.c_int32, ctypes.c_float, ctypes.c_float,
                                      ctypes.c_float, ctypes.c_float, ctypes.c_float, _fp]
    L.svh_simulate_rff.argtypes = [_vp, ctypes.c_int32, _i64p, _i32p, ctypes.c_int32, _dp, _dp,
                                   _dp, _dp, _dp]
    L.svh_autotune.argtypes = [_vp, ctypes.c_int32, ctypes.c_char_p]
    L.svh_autotune_read.argtypes = [_vp, ctypes.c_char_p, ctypes.c_int32]
    for name in ("svh_autotune", "svh_autotune_read", "svh_mlp_init", "svh_mlp_num_params", "svh_mlp_set_params", "svh_mlp_get_params",
                 "svh_mlp_forward", "svh_mlp_loss_grad", "svh_mlp_load_data",
                 "svh_mlp_train_epoch", "svh_simulate_rff"):
        getattr(L, name).restype = ctypes.c_int
    for name in ("svh_ordering_batch", "svh_batch_setup", "svh_batch_update", "svh_batch_data",
                 "svh_batch_rebuild", "svh_joint_solve", "svh_batch_fit_start", "svh_batch_fit_step",
                 "svh_batch_get", "svh_batch_set_reorder", "svh_batch_eval",
                 "svh_timing_enable", "svh_timing_read", "svh_timing_reset", "svh_maximin", "svh_ordered_knn", "svh_ordered_knn_rows", "svh_knn_query",
                 "svh_vecchia_setup", "svh_vecchia_rebuild", "svh_vecchia_eval", "svh_predict", "svh_predict_noalloc", "svh_block_columns",
                 "svh_create"):
        getattr(L, name).restype = ctypes.c_int
    _lib = L
    return L


EXPORTED = ("svh_device_count", "svh_version", "svh_create", "svh_destroy", "svh_last_error",
            "svh_maximin", "svh_ordered_knn", "svh_knn_query", "svh_vecchia_setup",
            "svh_vecchia_eval", "svh_predict", "svh_block_columns", "svh_timing_enable",
            "svh_timing_read", "svh_timing_reset", "svh_batch_setup", "svh_batch_update",
            "svh_batch_fit_start", "svh_batch_fit_step", "svh_batch_get", "svh_batch_set_reorder",
            "svh_batch_eval", "svh_ordering_batch", "svh_mlp_init", "svh_mlp_num_params",
            "svh_mlp_set_params", "svh_mlp_get_params", "svh_mlp_forward", "svh_mlp_loss_grad",
            "svh_mlp_load_data", "svh_mlp_train_epoch", "svh_simulate_rff", "svh_autotune",
            "svh_autotune_read", "svh_vecchia_rebuild", "svh_batch_data", "svh_batch_rebuild",
            "svh_joint_solve", "svh_ordered_knn_rows", "svh_predict_noalloc",
            "svh_matern_general")


def matern_general(nu, r):
    """Host evaluation of the general Matern M_nu(r) and g = M'/r (the
    kernels' code for smoothness outside {0.5, 1.5, 2.5})."""
    r = np.ascontiguousarray(r, dtype=np.float64).ravel()
    M, g = np.empty_like(r), np.empty_like(r)
    rc = _load().svh_matern_general(float(nu), _ptr(r), r.size, _ptr(M), _ptr(g))
    if rc != SVH_OK:
        raise ValueError("matern_general: nu must be positive")
    return M, g


def device_count() -> int:
    return int(_load().svh_device_count())


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _finite(a, what):
    a = _f64(a)
    if not np.all(np.isfinite(a)):
        raise ValueError(f"{what} contains NaN or infinite values")
    return a


def _ptr(a, t=_dp):
    return None if a is None else a.ctypes.data_as(t)


class Context:
    """One device context (one HIP stream).  Not thread-safe."""

    def __init__(self, device: int = 0):
        L = _load()
        h = _vp()
        rc = L.svh_create(int(device), ctypes.byref(h))
        if rc != SVH_OK:
            raise RuntimeError(f"svh_create(device={device}) failed with code {rc} "
                               "(no HIP device?)")
        self._h = h
        self._L = L
        self.device = device

    def close(self):
        if getattr(self, "_h", None):
            self._L.svh_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc == SVH_OK:
            return
        msg = self._L.svh_last_error(self._h).decode()
        if rc == SVH_ERR_CHOLESKY:
            raise np.linalg.LinAlgError(msg)
        if rc == SVH_ERR_ARG:
            raise ValueError(msg)
        raise RuntimeError(msg)

    # ---------------------------------------------------------------- timers
    def autotune(self, mode=2, cache_path=None):
        """Launch-schedule autotuner: 0 = built-in schedules, 1 = cached
        choices only, 2 = also time the candidates of new shape buckets.
        `cache_path`: text file of choices (read at first use, rewritten when
        a new bucket is tuned)."""
        path = None if cache_path is None else os.fsencode(cache_path)
        self._check(self._L.svh_autotune(self._h, int(mode), path))

    def autotune_choices(self):
        """{shape-bucket key: (candidate index, median ms)} tuned so far."""
        buf = ctypes.create_string_buffer(1 << 16)
        self._check(self._L.svh_autotune_read(self._h, buf, len(buf)))
        out = {}
        for line in buf.value.decode().splitlines():
            k, i, ms = line.split()
            out[k] = (int(i), float(ms))
        return out

    def timing(self, enable=True):
        self._check(self._L.svh_timing_enable(self._h, int(bool(enable))))

    def timing_reset(self):
        self._check(self._L.svh_timing_reset(self._h))

    def timing_read(self, name):
        ms = ctypes.c_double(0.0)
        cnt = ctypes.c_int64(0)
        self._check(self._L.svh_timing_read(self._h, name.encode(), ctypes.byref(ms),
                                            ctypes.byref(cnt)))
        return ms.value, cnt.value

    # ---------------------------------------------------------------- (a)
    def maximin(self, X, first):
        X = _finite(X, "X")
        n, d = X.shape
        out = np.empty(n, dtype=np.int64)
        self._check(self._L.svh_maximin(self._h, _ptr(X), n, d, int(first), _ptr(out, _i64p)))
        return out

    def ordered_knn(self, X, m, start=0):
        X = _finite(X, "X")
        n, d = X.shape
        nn = np.empty((n, m + 1), dtype=np.int64)
        self._check(self._L.svh_ordered_knn(self._h, _ptr(X), n, d, int(m), int(start),
                                            _ptr(nn, _i64p)))
        return nn

    def ordered_knn_rows(self, X, m, start):
        """Rows start..n-1 of ordered_knn(X, m, start)."""
        X = _finite(X, "X")
        n, d = X.shape
        nn = np.empty((n - int(start), m + 1), dtype=np.int64)
        self._check(self._L.svh_ordered_knn_rows(self._h, _ptr(X), n, d, int(m), int(start),
                                                 _ptr(nn, _i64p)))
        return nn

    def knn_query(self, Xref, Xq, k):
        Xref, Xq = _finite(Xref, "Xref"), _finite(Xq, "Xq")
        nb = np.empty((Xq.shape[0], k), dtype=np.int64)
        self._check(self._L.svh_knn_query(self._h, _ptr(Xref), Xref.shape[0], _ptr(Xq),
                                          Xq.shape[0], Xref.shape[1], int(k), _ptr(nb, _i64p)))
        return nb

    def ordering_batch(self, Xs, m):
        """Maximin orders and ordered m-NN of several datasets in one pass."""
        Xs = [_finite(X, "X") for X in Xs]
        n = np.array([X.shape[0] for X in Xs], np.int64)
        d = np.array([X.shape[1] for X in Xs], np.int32)
        from .ordering import maximin_first
        first = np.array([maximin_first(X) for X in Xs], np.int64)
        Xc = np.concatenate([X.ravel() for X in Xs])
        order = np.empty(int(n.sum()), np.int64)
        nn = np.empty((int(n.sum()), m + 1), np.int64)
        self._check(self._L.svh_ordering_batch(self._h, len(Xs), _ptr(n, _i64p), _ptr(d, _i32p),
                                               _ptr(first, _i64p), _ptr(Xc), int(m),
                                               _ptr(order, _i64p), _ptr(nn, _i64p)))
        out, o = [], 0
        for k in n:
            out.append((order[o:o + k], nn[o:o + k]))
            o += k
        return out

    # ---------------------------------------------------------------- (b)
    def vecchia_setup(self, Xo, yo, Zo, rows, nbr, kcnt, mmax, chunk_grad, chunk_val,
                      nchunk_grad, nchunk_val):
        Xo, yo = _finite(Xo, "Xo"), _finite(yo, "yo")
        q = 0 if Zo is None else Zo.shape[1]
        Zo = None if q == 0 else _f64(Zo)
        rows = np.ascontiguousarray(rows, dtype=np.int64)
        nbr = np.ascontiguousarray(nbr, dtype=np.int64)
        kcnt = np.ascontiguousarray(kcnt, dtype=np.int32)
        cg = np.ascontiguousarray(chunk_grad, dtype=np.int32)
        cv = np.ascontiguousarray(chunk_val, dtype=np.int32)
        self._check(self._L.svh_vecchia_setup(
            self._h, _ptr(Xo), _ptr(yo), _ptr(Zo), Xo.shape[0], Xo.shape[1], q, rows.shape[0],
            _ptr(rows, _i64p), _ptr(nbr, _i64p), _ptr(kcnt, _i32p), int(mmax), _ptr(cg, _i32p),
            _ptr(cv, _i32p), int(nchunk_grad), int(nchunk_val)))
        self._q = q

    def vecchia_rebuild(self, Xe, ye, Ze, lam, first, m, upload=True):
        """Device rebuild at ranges `lam` (svh_vecchia_rebuild); returns the
        maximin order.  With upload=False the data of the previous call are
        reused."""
        lam = _f64(lam)
        n, d = (Xe.shape if Xe is not None else (self._rb_n, self._rb_d))
        if upload:
            Xe, ye = _finite(Xe, "Xe"), _finite(ye, "ye")
            q = 0 if Ze is None else Ze.shape[1]
            Ze = None if q == 0 else _f64(Ze)
            self._rb_n, self._rb_d, self._rb_q = n, d, q
        q = self._rb_q
        order = np.empty(n, dtype=np.int64)
        self._check(self._L.svh_vecchia_rebuild(
            self._h, _ptr(Xe) if upload else None, _ptr(ye) if upload else None,
            _ptr(Ze) if (upload and q) else None, n, d, q, _ptr(lam), int(first), int(m),
            1 if upload else 0, _ptr(order, _i64p)))
        self._q = q
        return order

    def _eval_buffers(self, P):
        """One host array per (P, q) for an evaluation's input and outputs,
        [theta | ll | grad | Fisher | beta], with its ctypes pointers made once
        (building pointer objects costs more than the copies per call)."""
        q = self._q
        key = (P, q)
        cache = getattr(self, "_evbuf", None)
        if cache is None:
            cache = self._evbuf = {}
        ent = cache.get(key)
        if ent is None:
            arr = np.zeros(P + 1 + P + P * P + max(q, 1))
            base = arr.ctypes.data
            offs = (0, P, P + 1, 2 * P + 1, 2 * P + 1 + P * P)
            ptrs = tuple(ctypes.cast(base + 8 * o, _dp) for o in offs)
            jp = ctypes.c_int32(0)
            ent = cache[key] = (arr, offs, ptrs, jp, ctypes.byref(jp))
        return ent

    def vecchia_eval(self, theta, nu, need_grad=True, var_penalty=0.0, log_var_target=0.0,
                     precision="fp64"):
        theta = _f64(theta)
        if theta.ndim != 1:
            raise ValueError("theta must be 1-D")
        P = theta.shape[0]
        q = self._q
        arr, offs, ptrs, jp, jref = self._eval_buffers(P)
        arr[:P] = theta
        self._check(self._L.svh_vecchia_eval(
            self._h, ptrs[0], P, float(nu), PRECISION[precision], int(bool(need_grad)),
            float(var_penalty), float(log_var_target), ptrs[1], ptrs[2], ptrs[3],
            ptrs[4] if q else None, jref))
        self.last_jitter_power = jp.value
        ll = float(arr[offs[1]])
        beta = arr[offs[4]:offs[4] + q].copy()
        if not need_grad:
            return ll, None, None, beta
        grad = arr[offs[2]:offs[2] + P].copy()
        fish = arr[offs[3]:offs[3] + P * P].reshape(P, P).copy()
        return ll, grad, fish, beta

    # ---------------------------------------------------------------- (c)
    def predict(self, Xtr, resid, Xq, m, theta, nu, noise_free=True, precision="fp64",
                return_nb=False):
        Xtr, resid, Xq = _finite(Xtr, "X"), _finite(resid, "resid"), _finite(Xq, "Xstar")
        theta = _finite(theta, "theta")
        nq = Xq.shape[0]
        mean = np.empty(nq)
        var = np.empty(nq)
        nb = np.empty((nq, m), dtype=np.int64) if return_nb else None
        self._check(self._L.svh_predict(
            self._h, _ptr(Xtr), _ptr(resid), Xtr.shape[0], _ptr(Xq), nq, Xtr.shape[1], int(m),
            _ptr(theta), float(nu), PRECISION[precision], int(bool(noise_free)), _ptr(mean),
            _ptr(var), _ptr(nb, _i64p)))
        return (mean, var, nb) if return_nb else (mean, var)

    def predict_noalloc(self, Xtr, resid, ntr, Xq, nq, d, m, theta, nu, mean, var,
                        noise_free=True, precision="fp64"):
        """svh_predict_noalloc: Xtr, resid, Xq, mean, var are DEVICE addresses
        (ints) on this context's GPU; theta is a host array."""
        theta = _finite(theta, "theta")
        self._check(self._L.svh_predict_noalloc(
            self._h, Xtr, resid, int(ntr), Xq, int(nq), int(d), int(m), _ptr(theta), float(nu),
            PRECISION[precision], int(bool(noise_free)), mean, var))

    def block_columns(self, Xall, blocks, mask, theta, nu, precision="fp64"):
        Xall, theta = _finite(Xall, "X"), _finite(theta, "theta")
        blocks = np.ascontiguousarray(blocks, dtype=np.int64)
        mask = None if mask is None else _f64(mask)
        nblk, K = blocks.shape
        a = np.empty((nblk, K))
        self._check(self._L.svh_block_columns(
            self._h, _ptr(Xall), Xall.shape[0], Xall.shape[1], _ptr(blocks, _i64p), _ptr(mask),
            nblk, K, _ptr(theta), float(nu), PRECISION[precision], _ptr(a)))
        return a

    def joint_solve(self, nobs, blocks, a, resid=None, rhs=None, exact_var=False):
        """Solves with Lt = U_pp^T of the joint factor (svh_joint_solve).
        Returns (mean or None, Lt^-1 rhs or None, exact variances or None)."""
        blocks = np.ascontiguousarray(blocks, dtype=np.int64)
        a = _f64(a)
        ns, K = blocks.shape
        mean = np.empty(ns) if resid is not None else None
        resid = None if resid is None else _f64(resid)
        nrhs = 0
        x = None
        if rhs is not None:
            rhs = _f64(rhs).reshape(ns, -1)
            nrhs = rhs.shape[1]
            x = np.empty((ns, nrhs))
        var = np.empty(ns) if exact_var else None
        self._check(self._L.svh_joint_solve(
            self._h, int(nobs), ns, K, _ptr(blocks, _i64p), _ptr(a), _ptr(resid), nrhs,
            _ptr(rhs), _ptr(mean), _ptr(x), 1 if exact_var else 0, _ptr(var)))
        return mean, x, var


BP, BQ = 12, 11   # parameter / trend strides of the batched entry points


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


_fptr = ctypes.POINTER(ctypes.c_float)


class MlpHandle:
    """The context's warm-start network (svh_mlp_*): bf16-weight matrix-core
    forward/backward and Adam on the device."""

    def __init__(self, ctx: "Context", n_in, n_hidden, n_layers, n_out, max_batch=1024):
        self.ctx = ctx
        self.shape = (n_in, n_hidden, n_layers, n_out)
        ctx._check(ctx._L.svh_mlp_init(ctx._h, n_in, n_hidden, n_layers, n_out, max_batch))
        n = ctypes.c_int64()
        ctx._check(ctx._L.svh_mlp_num_params(ctx._h, ctypes.byref(n)))
        self.n_params = int(n.value)

    def set_params(self, p):
        p = _f32(p)
        if p.size != self.n_params:
            raise ValueError(f"expected {self.n_params} parameters, got {p.size}")
        self.ctx._check(self.ctx._L.svh_mlp_set_params(self.ctx._h, p.ctypes.data_as(_fptr)))

    def get_params(self):
        p = np.empty(self.n_params, np.float32)
        self.ctx._check(self.ctx._L.svh_mlp_get_params(self.ctx._h, p.ctypes.data_as(_fptr)))
        return p

    def forward(self, X):
        X = _f32(np.atleast_2d(X))
        if X.shape[1] != self.shape[0]:
            raise ValueError("forward: wrong input width")
        out = np.empty((X.shape[0], self.shape[3]), np.float32)
        self.ctx._check(self.ctx._L.svh_mlp_forward(self.ctx._h, X.ctypes.data_as(_fptr),
                                                    X.shape[0], out.ctypes.data_as(_fptr)))
        return out

    def loss_grad(self, X, T, M):
        X, T, M = _f32(X), _f32(T), _f32(M)
        loss = ctypes.c_float()
        g = np.empty(self.n_params, np.float32)
        self.ctx._check(self.ctx._L.svh_mlp_loss_grad(
            self.ctx._h, X.ctypes.data_as(_fptr), T.ctypes.data_as(_fptr),
            M.ctypes.data_as(_fptr), X.shape[0], ctypes.byref(loss), g.ctypes.data_as(_fptr)))
        return float(loss.value), g

    def load_data(self, X, T, M):
        X, T, M = _f32(X), _f32(T), _f32(M)
        self._n = X.shape[0]
        self.ctx._check(self.ctx._L.svh_mlp_load_data(
            self.ctx._h, X.ctypes.data_as(_fptr), T.ctypes.data_as(_fptr),
            M.ctypes.data_as(_fptr), X.shape[0]))

    def train_epoch(self, perm, batch, lr, beta1=0.9, beta2=0.999, eps=1e-8, weight_decay=0.0):
        perm = np.ascontiguousarray(perm, dtype=np.int32)
        if perm.size != self._n or perm.min() < 0 or perm.max() >= self._n:
            raise ValueError("train_epoch: perm must index the loaded rows")
        loss = ctypes.c_float()
        self.ctx._check(self.ctx._L.svh_mlp_train_epoch(
            self.ctx._h, perm.ctypes.data_as(_i32p), int(batch), float(lr), float(beta1),
            float(beta2), float(eps), float(weight_decay), ctypes.byref(loss)))
        return float(loss.value)


def simulate_rff(ctx: "Context", Xs, Ws, bs, amps):
    """Random-Fourier-feature draws y = amp * sum_f cos(W_f . x + b_f) for
    several datasets (coordinates already divided by the ranges)."""
    nd = len(Xs)
    n = np.array([X.shape[0] for X in Xs], np.int64)
    d = np.array([X.shape[1] for X in Xs], np.int32)
    F = Ws[0].shape[0]
    X = np.concatenate([_f64(x).ravel() for x in Xs])
    W = np.concatenate([_f64(w).ravel() for w in Ws])
    b = np.concatenate([_f64(v) for v in bs])
    amp = _f64(amps)
    y = np.empty(int(n.sum()))
    ctx._check(ctx._L.svh_simulate_rff(ctx._h, nd, n.ctypes.data_as(_i64p),
                                       d.ctypes.data_as(_i32p), F, _ptr(X), _ptr(W), _ptr(b),
                                       _ptr(amp), _ptr(y)))
    return np.split(y, np.cumsum(n)[:-1])


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


class BatchHandle:
    """Thin wrapper of the svh_batch_* entry points on one context."""

    def __init__(self, ctx: "Context"):
        self.ctx = ctx
        self._L = ctx._L

    def setup(self, d, q, nu_code, n, X, y, Z, nbr, kcnt, chunk, nchunk, mmax, lvt, fixed_nug,
              log_nug, lambda_max):
        """X, y, Z and nbr, kcnt may be None: buffers only, filled by
        data() + rebuild()."""
        self.n = _i64(n)
        self.ndata = len(self.n)
        if X is not None:
            X, y = _finite(X, "X"), _finite(y, "y")
        Z = _f64(Z) if Z is not None and Z.size else np.zeros(1)
        nbr = None if nbr is None else _i64(nbr)
        kcnt = None if kcnt is None else _i32(kcnt)
        args = [_i32(d), _i32(q), _i32(nu_code), self.n]
        self._keep = args
        c = self.ctx
        c._check(self._L.svh_batch_setup(
            c._h, self.ndata, _ptr(args[0], _i32p), _ptr(args[1], _i32p), _ptr(args[2], _i32p),
            _ptr(args[3], _i64p), _ptr(X), _ptr(y), _ptr(Z), _ptr(nbr, _i64p),
            _ptr(kcnt, _i32p), _ptr(_i32(chunk), _i32p), _ptr(_i32(nchunk), _i32p),
            int(mmax), _ptr(_f64(lvt)), _ptr(_i32(fixed_nug), _i32p), _ptr(_f64(log_nug)),
            float(lambda_max)))

    def update(self, i, X, y, Z, nbr, kcnt):
        Z = _f64(Z) if Z is not None and Z.size else np.zeros(1)
        nbr, kcnt = _i64(nbr), _i32(kcnt)
        c = self.ctx
        c._check(self._L.svh_batch_update(c._h, int(i), _ptr(_finite(X, "X")),
                                          _ptr(_finite(y, "y")), _ptr(Z), _ptr(nbr, _i64p),
                                          _ptr(kcnt, _i32p)))

    def data(self, Xe, ye, Ze):
        """Unordered estimation data of every dataset (setup layout), kept on
        the device for `rebuild`."""
        self._data = (_finite(Xe, "Xe"), _finite(ye, "ye"),
                      _f64(Ze) if Ze is not None and Ze.size else np.zeros(1))
        c = self.ctx
        c._check(self._L.svh_batch_data(c._h, _ptr(self._data[0]), _ptr(self._data[1]),
                                        _ptr(self._data[2])))

    def rebuild(self, which, lams, firsts, ms):
        """Device rebuild of datasets `which` at ranges lams[k] (svh_batch_rebuild)."""
        which = _i32(which)
        k = len(which)
        lam = np.ones((k, 12))
        for r, la in enumerate(lams):
            lam[r, :len(la)] = la
        c = self.ctx
        c._check(self._L.svh_batch_rebuild(c._h, k, _ptr(which, _i32p), _ptr(lam),
                                           _ptr(_i64(firsts), _i64p), _ptr(_i32(ms), _i32p)))

    def fit_start(self, theta0, max_iter, tol, var_penalty, use_graph=True):
        th = _f64(theta0)
        c = self.ctx
        c._check(self._L.svh_batch_fit_start(c._h, _ptr(th), int(max_iter), float(tol),
                                             float(var_penalty), int(bool(use_graph))))

    def fit_step(self, nrep=1):
        st = np.zeros(self.ndata, dtype=np.int32)
        c = self.ctx
        c._check(self._L.svh_batch_fit_step(c._h, int(nrep), _ptr(st, _i32p)))
        return st

    def get(self):
        nd = self.ndata
        th = np.zeros((nd, BP))
        ll = np.zeros(nd)
        beta = np.zeros((nd, BQ))
        st = np.zeros(nd, np.int32)
        it = np.zeros(nd, np.int32)
        ne = np.zeros(nd, np.int32)
        c = self.ctx
        c._check(self._L.svh_batch_get(c._h, _ptr(th), _ptr(ll), _ptr(beta), _ptr(st, _i32p),
                                       _ptr(it, _i32p), _ptr(ne, _i32p)))
        return th, ll, beta, st, it, ne

    def set_reorder(self, i, rebuilt):
        c = self.ctx
        c._check(self._L.svh_batch_set_reorder(c._h, int(i), int(bool(rebuilt))))

    def eval(self, theta):
        nd = self.ndata
        th = _f64(theta)
        ll = np.zeros(nd)
        g = np.zeros((nd, BP))
        F = np.zeros((nd, BP * BP))
        beta = np.zeros((nd, BQ))
        jr = ctypes.c_int32(0)
        c = self.ctx
        c._check(self._L.svh_batch_eval(c._h, _ptr(th), _ptr(ll), _ptr(g), _ptr(F), _ptr(beta),
                                        ctypes.byref(jr)))
        return ll, g, F, beta, jr.value


def context(device: int = 0) -> Context:
    """Process-wide default context on `device` (created on first use)."""
    global _ctx
    with _lock:
        if _ctx is None or _ctx.device != device:
            _ctx = Context(device)
        return _ctx
