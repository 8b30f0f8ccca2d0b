"""Shared test inputs."""

from fractions import Fraction

import numpy as np


def _seq_sq(t):
    s = 0.0
    for v in t:
        s = s + v * v
    return s


def _fused_sq(t, k):
    """2-D squared norm with term k's product fused into the add (one FMA),
    evaluated exactly and rounded once."""
    o = 1 - k
    return float(Fraction(float(t[k])) ** 2 + Fraction(float(t[o] * t[o])))


def near_tie_pairs(count=24, seed=0):
    """Queries with two candidates each, where the canonical squared distance
    (sequential sum of rounded products) makes the second candidate strictly
    nearer but either FMA-contracted evaluation makes the first one nearer or
    equal (so the lower index would win).  Groups sit 8 apart, so each query's
    nearest reference is one of its own pair.

    Returns (Xref (2*count, 2), Xq (count, 2)); the canonical nearest
    reference of query k is 2k + 1."""
    rng = np.random.default_rng(seed)
    refs, qs = [], []
    while len(qs) < count:
        k = len(qs)
        x = np.array([8.0 * k + 0.25, 0.5])
        y1 = x + rng.uniform(0.3, 0.9, 2)
        t1 = y1 - x
        s1 = _seq_sq(t1)
        a = rng.uniform(0.3, 0.9)
        r = s1 - a * a
        if r <= 0:
            continue
        y2 = x + np.array([a, np.sqrt(r)])
        t2 = y2 - x
        s2 = _seq_sq(t2)
        if s1 == s2:
            continue
        (w, tw), (lo, tl) = ((y1, t1), (y2, t2)) if s1 < s2 else ((y2, t2), (y1, t1))
        if all(_fused_sq(tl, j) <= _fused_sq(tw, j) for j in (0, 1)):
            refs += [lo, w]
            qs.append(x)
    return np.array(refs), np.array(qs)
