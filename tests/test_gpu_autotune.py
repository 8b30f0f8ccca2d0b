"""Launch-schedule autotuner: tuned schedules give the same results as the
built-in ones, and the choices persist through the cache file."""

import math

import numpy as np
import pytest

from oracle import vecchia_np as O

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def test_tuned_schedules_match_defaults(gpu, tmp_path):
    from pyscaledvecchia_amd import VecchiaStructure
    rng = np.random.default_rng(5)
    n, d, m = 4000, 5, 30
    X = rng.random((n, d))
    y = np.sin(4 * X[:, 0]) + X[:, 1] + 0.01 * rng.standard_normal(n)
    theta = np.concatenate([[0.0], np.full(d, math.log(0.3)), [math.log(1e-3)]])
    Xs = X / np.exp(theta[1:1 + d])
    order = O.maximin_order(Xs)
    groups = O.nn_groups(O.find_ordered_nn(Xs[order], m), m)
    Z = np.column_stack([np.ones(n), X])[order]
    st = VecchiaStructure(X[order], y[order], Z, groups, ctx=gpu)
    Xt = rng.random((300, d))
    try:
        gpu.autotune(0)
        ref = st(theta, 2.5)
        mu0, va0 = gpu.predict(X, y, Xt, 140, theta, 2.5)
        cache = str(tmp_path / "tune.txt")
        gpu.autotune(2, cache)
        res = st(theta, 2.5)
        mu, va = gpu.predict(X, y, Xt, 140, theta, 2.5)
        ch = gpu.autotune_choices()
        assert any(k.startswith("ll_pass1:") for k in ch)
        assert any(k.startswith("pred_blocks:") for k in ch)
        # same arithmetic per block; only the grouping of the partial sums moves
        assert abs(res[0] - ref[0]) <= 1e-12 * abs(ref[0])
        for a, b in zip(res[1:], ref[1:]):
            assert _rel(a, b) < 1e-10
        assert _rel(mu, mu0) < 1e-12 and _rel(va, va0) < 1e-12
        # a fresh reader of the cache gets the same choices without timing
        gpu.autotune(1, cache)
        st(theta, 2.5)
        assert gpu.autotune_choices() == ch
    finally:
        gpu.autotune(0)
