"""Maximin ordering and ordered nearest neighbours on the GPU.

Drop-in for scaled_vecchia.ordering (reference ordering.py:62-127): same
names, arguments and output layout; the O(n^2 d) work runs in libsvh.so.
"""

from __future__ import annotations

import numpy as np

from . import _lib

__all__ = ["maximin_order", "find_ordered_nn"]


def maximin_first(X: np.ndarray, rows: int = 8192) -> int:
    """The point closest to the centroid (ordering.py:72):
    argmin(((X - X.mean(0)) ** 2).sum(1)), first index on ties.  The same
    NumPy operations run on blocks of rows (each row's sum is independent of
    the others), so the result is that of the one-pass expression without its
    full-size temporaries."""
    mu = X.mean(0)
    best, at = None, -1
    for s in range(0, X.shape[0], rows):
        v = ((X[s:s + rows] - mu) ** 2).sum(1)
        i = int(np.argmin(v))
        if best is None or v[i] < best:
            best, at = v[i], s + i
    return at


def maximin_order(X: np.ndarray, first: int | None = None) -> np.ndarray:
    """Exact maximum-minimum-distance ordering (ordering.py:62-73).

    The first point is the one closest to the centroid, chosen on the host
    exactly as ordering.py:72 does; the n-1 farthest-point steps run on the
    GPU (lowest index wins ties).
    """
    X = np.ascontiguousarray(X, dtype=np.float64)
    if X.ndim != 2:
        raise ValueError("X must be 2-D")
    if X.shape[0] == 0:
        return np.empty(0, dtype=np.int64)
    if first is None:
        first = maximin_first(X)
    return _lib.context().maximin(X, int(first))


def find_ordered_nn(X: np.ndarray, m: int, start: int = 0,
                    block_size: int = 1024) -> np.ndarray:
    """For each i, the min(i, m) nearest neighbours among indices < i
    (ordering.py:76-116).  Column 0 is i, -1 pads; neighbours are sorted
    nearest first (ties by index).  `block_size` is accepted for signature
    compatibility and unused.
    """
    X = np.ascontiguousarray(X, dtype=np.float64)
    return _lib.context().ordered_knn(X, int(m), int(start))


def _nn_groups(nn: np.ndarray, m: int):
    """Rows split by conditioning-set size k = min(i, m) (ordering.py:119-127)."""
    n = nn.shape[0]
    ks = np.minimum(np.arange(n), m)
    out = []
    for k in np.unique(ks):
        rows = np.flatnonzero(ks == k)
        out.append((rows, nn[rows, 1:k + 1]))
    return out
