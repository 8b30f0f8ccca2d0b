"""The C restatement of the oracle agrees with the NumPy oracle (CPU)."""

import math

import numpy as np
import pytest

from oracle import c_oracle as C
from oracle import vecchia_np as O

pytestmark = pytest.mark.skipif(not C.available(), reason="liboracle_c.so not built")


@pytest.mark.parametrize("n,d,seed", [(300, 3, 0), (1000, 8, 1)])
def test_c_maximin_equals_numpy(n, d, seed):
    X = np.random.default_rng(seed).random((n, d))
    f = O.maximin_first(X)
    np.testing.assert_array_equal(C.maximin_order(X, f), O.maximin_order(X, f))


def test_c_knn_equals_numpy():
    X = np.random.default_rng(3).random((800, 4))
    for start in (0, 500):
        np.testing.assert_array_equal(C.find_ordered_nn(X, 30, start),
                                      O.find_ordered_nn(X, 30, start))


@pytest.mark.parametrize("nthreads", [1, 4])
def test_c_loglik_equals_numpy(nthreads):
    rng = np.random.default_rng(4)
    n, d = 1500, 5
    X = rng.random((n, d))
    y = np.sin(5 * X[:, 0]) + X[:, 1] + 0.01 * rng.standard_normal(n)
    theta = np.log([1.2, 0.3, 0.5, 0.8, 2.0, 5.0, 1e-3])
    Xs = X / np.exp(theta[1:1 + d])
    o = O.maximin_order(Xs)
    groups = O.nn_groups(O.find_ordered_nn(Xs[o], 20), 20)
    Z = np.column_stack([np.ones(n), X])[o]
    a = C.vecchia_profile_loglik(theta, X[o], y[o], Z, groups, 1.5, 1.0, 0.2, nthreads)
    b = O.vecchia_profile_loglik(theta, X[o], y[o], Z, groups, 1.5, var_penalty=1.0,
                                 log_var_target=0.2)
    assert abs(a[0] - b[0]) < 1e-10 * abs(b[0])
    for u, v in zip(a[1:], b[1:]):
        assert np.max(np.abs(u - v)) < 1e-9 * max(1.0, np.max(np.abs(v)))
