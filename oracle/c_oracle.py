"""ctypes wrapper of liboracle_c.so (C restatement; test infrastructure only)."""

from __future__ import annotations

import ctypes
import os

import numpy as np

_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "liboracle_c.so")
_L = None
_dp = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)
_i32p = ctypes.POINTER(ctypes.c_int32)


def available():
    return os.path.exists(_PATH)


def _lib():
    global _L
    if _L is None:
        L = ctypes.CDLL(_PATH)
        L.oracle_maximin.argtypes = [_dp, ctypes.c_int64, ctypes.c_int, ctypes.c_int64, _i64p]
        L.oracle_maximin.restype = None
        L.oracle_ordered_knn.argtypes = [_dp, ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int64, _i64p]
        L.oracle_ordered_knn.restype = None
        L.oracle_vecchia_eval.argtypes = [_dp, _dp, _dp, ctypes.c_int64, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_int64, _i64p, _i64p, _i32p,
                                          ctypes.c_int, _dp, ctypes.c_int, ctypes.c_double,
                                          ctypes.c_double, ctypes.c_int, _dp, _dp, _dp, _dp]
        L.oracle_vecchia_eval.restype = ctypes.c_int
        _L = L
    return _L


def _p(a, t=_dp):
    return a.ctypes.data_as(t)


def maximin_order(X, first):
    X = np.ascontiguousarray(X, dtype=np.float64)
    out = np.empty(X.shape[0], dtype=np.int64)
    _lib().oracle_maximin(_p(X), X.shape[0], X.shape[1], int(first), _p(out, _i64p))
    return out


def find_ordered_nn(X, m, start=0):
    X = np.ascontiguousarray(X, dtype=np.float64)
    nn = np.empty((X.shape[0], m + 1), dtype=np.int64)
    _lib().oracle_ordered_knn(_p(X), X.shape[0], X.shape[1], int(m), int(start), _p(nn, _i64p))
    return nn


def blocks_from_groups(groups):
    mmax = max(nb.shape[1] for _, nb in groups)
    nblk = sum(len(r) for r, _ in groups)
    rows = np.empty(nblk, np.int64)
    nbr = np.full((nblk, max(mmax, 1)), -1, np.int64)
    kc = np.empty(nblk, np.int32)
    o = 0
    for r, nb in groups:
        rows[o:o + len(r)] = r
        nbr[o:o + len(r), :nb.shape[1]] = nb
        kc[o:o + len(r)] = nb.shape[1]
        o += len(r)
    return rows, nbr, kc, max(mmax, 1)


def vecchia_profile_loglik(theta, Xo, yo, Zo, groups, nu, var_penalty=0.0,
                           log_var_target=0.0, nthreads=1):
    Xo = np.ascontiguousarray(Xo, np.float64)
    yo = np.ascontiguousarray(yo, np.float64)
    n, d = Xo.shape
    q = 0 if Zo is None else Zo.shape[1]
    Zo = np.zeros((n, 1)) if q == 0 else np.ascontiguousarray(Zo, np.float64)
    rows, nbr, kc, mmax = blocks_from_groups(groups)
    theta = np.ascontiguousarray(theta, np.float64)
    P = d + 2
    ll = np.zeros(1)
    g = np.zeros(P)
    F = np.zeros((P, P))
    beta = np.zeros(max(q, 1))
    code = {0.5: 0, 1.5: 1, 2.5: 2}[nu]
    rc = _lib().oracle_vecchia_eval(_p(Xo), _p(yo), _p(Zo), n, d, q, len(rows), _p(rows, _i64p),
                                    _p(nbr, _i64p), _p(kc, _i32p), mmax, _p(theta), code,
                                    float(var_penalty), float(log_var_target), int(nthreads),
                                    _p(ll), _p(g), _p(F), _p(beta))
    if rc:
        raise np.linalg.LinAlgError("block Cholesky failed")
    return float(ll[0]), g, F, beta[:q]
