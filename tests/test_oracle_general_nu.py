"""General and estimated Matern smoothness (reference _matern_general.py and
the estimated-nu paths of likelihood.py / gp.py), CPU only.

Pins of the oracle restate the reference's own tests
(tests/test_likelihood.py:104-189, tests/test_gp.py:134-182): the Bessel form
reduces to the closed forms at nu = 0.5, 1.5, 2.5; g = M'/r against a finite
difference at nu = 1.75; the estimated-nu gradient (including d/dlog nu,
itself a central difference in nu) against finite differences of the
log-likelihood; Vecchia equals the exact GP at m = n - 1 with nu estimated.
The library's host evaluation of the kernels' Bessel code
(svh_matern_general) is checked against the oracle (scipy.special.kv).
"""

import math

import numpy as np
import pytest
from scipy.spatial.distance import cdist

from oracle import vecchia_np as O


def test_general_matern_reduces_to_closed_forms():
    r = np.linspace(1e-6, 5, 200)
    for nu in (0.5, 1.5, 2.5):
        c = math.sqrt(2 * nu) * r
        closed = {0.5: np.exp(-c), 1.5: (1 + c) * np.exp(-c),
                  2.5: (1 + c + c * c / 3) * np.exp(-c)}[nu]
        np.testing.assert_allclose(O.matern_corr_general(r, nu), closed, atol=1e-10)
        mask = r > 0.01
        gc = {0.5: -np.exp(-r) / r, 1.5: -3 * np.exp(-c),
              2.5: -(5 / 3) * (1 + c) * np.exp(-c)}[nu]
        np.testing.assert_allclose(O.matern_g_general(r, nu)[mask], gc[mask], atol=1e-8)


def test_general_matern_g_matches_numerical_derivative():
    r = np.linspace(0.05, 5, 200)
    nu, h = 1.75, 1e-6
    g_num = (O.matern_corr_general(r + h, nu) - O.matern_corr_general(r - h, nu)) / (2 * h) / r
    assert np.max(np.abs(g_num - O.matern_g_general(r, nu)) / np.abs(g_num)) < 1e-6


def _structure(X, theta, d, m):
    Xs = X / np.exp(theta[1:1 + d])
    order = O.maximin_order(Xs)
    nn = O.find_ordered_nn(Xs[order], m)
    return order, O.nn_groups(nn, m)


def test_estimate_nu_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    n, d, m = 60, 3, 10
    X = rng.random((n, d))
    y = rng.standard_normal(n)
    Z = np.ones((n, 1))
    theta = np.array([math.log(1.7), math.log(0.35), math.log(0.9), math.log(0.5),
                      math.log(1.8), math.log(0.05)])
    order, groups = _structure(X, theta, d, m)

    def f(th):
        return O.vecchia_profile_loglik(th, X[order], y[order], Z[order], groups, None,
                                        estimate_nu=True)
    ll0, g0, M0, _ = f(theta)
    num = np.empty_like(g0)
    h = 1e-6
    for j in range(len(theta)):
        tp, tm = theta.copy(), theta.copy()
        tp[j] += h
        tm[j] -= h
        num[j] = (f(tp)[0] - f(tm)[0]) / (2 * h)
    assert np.max(np.abs(g0 - num) / (1 + np.abs(num))) < 1e-5
    assert np.abs(M0 - M0.T).max() < 1e-6
    assert np.linalg.eigvalsh(0.5 * (M0 + M0.T)).min() > -1e-6


def test_estimate_nu_matches_exact_gp_at_full_conditioning():
    rng = np.random.default_rng(2)
    n, d = 50, 2
    X = rng.random((n, d))
    y = rng.standard_normal(n)
    theta = np.array([math.log(1.2), math.log(0.4), math.log(0.3), math.log(1.75),
                      math.log(0.02)])
    var, lam = math.exp(theta[0]), np.exp(theta[1:1 + d])
    nu, nug = math.exp(theta[1 + d]), math.exp(theta[2 + d])
    K = var * (O.matern_corr_general(cdist(X / lam, X / lam), nu) + nug * np.eye(n))
    L = np.linalg.cholesky(K)
    a = np.linalg.solve(L, y)
    exact = -0.5 * (n * O.LOG2PI + 2 * np.log(np.diag(L)).sum() + a @ a)
    order, groups = _structure(X, theta, d, n - 1)
    ll = O.vecchia_profile_loglik(theta, X[order], y[order], None, groups, None,
                                  estimate_nu=True)[0]
    assert abs(ll - exact) < 1e-7


def test_oracle_fit_with_fixed_general_and_estimated_nu():
    rng = np.random.default_rng(20)
    X = rng.random((150, 2))
    y = np.sin(3 * X[:, 0]) + 0.1 * X[:, 1]
    g = O.OracleGP(m_est=10, m_pred=20, n_est=150, nu=1.75).fit(X, y)
    assert g.nu_ == pytest.approx(1.75)
    assert np.all(np.isfinite(g.predict(X[:5])))
    g2 = O.OracleGP(m_est=10, m_pred=20, n_est=150, nu=None, var_correction=False).fit(X, y)
    assert g2._estimate_nu and 0.3 <= g2.nu_ <= 10.0 * (1 + 1e-12) and len(g2.theta_) == 5


def test_library_host_matern_matches_oracle():
    """svh_matern_general (the kernels' Temme / continued-fraction Bessel code,
    run on the host) against scipy.special.kv through the oracle."""
    from pyscaledvecchia_amd import _lib
    r = np.concatenate([[0.0, 5e-8], np.logspace(-6, 1.8, 300)])
    for nu in (0.3, 0.45, 0.5, 0.75, 1.0, 1.5, 1.75, 2.5, 3.2, 5.0, 7.6, 10.0):
        M, g = _lib.matern_general(nu, r)
        M0, g0 = O.matern_corr_general(r, nu), O.matern_g_general(r, nu)
        ok = M0 > 1e-250
        np.testing.assert_allclose(M[ok], M0[ok], rtol=1e-12, atol=0)
        np.testing.assert_allclose(g[ok], g0[ok], rtol=1e-11, atol=0)
    with pytest.raises(ValueError):
        _lib.matern_general(0.0, r)
