"""General and estimated Matern smoothness on the GPU (Bessel K_nu in the
kernels) against the oracle (scipy.special.kv): fixed non-half-integer nu,
estimated nu (theta = [log sigma^2, log lambda, log nu, log tau]), the exact
GP at full conditioning, whole fits and predictions."""

import math

import numpy as np
import pytest
from scipy.spatial.distance import cdist

from oracle import vecchia_np as O

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def _structure(X, theta, d, m):
    Xs = X / np.exp(theta[1:1 + d])
    order = O.maximin_order(Xs)
    nn = O.find_ordered_nn(Xs[order], m)
    return order, O.nn_groups(nn, m)


@pytest.mark.parametrize("nu", [0.35, 0.8, 1.75, 4.2])
@pytest.mark.parametrize("trend", ["zero", "linear"])
def test_fixed_general_nu_loglik(gpu, nu, trend):
    from pyscaledvecchia_amd.likelihood import vecchia_profile_loglik
    rng = np.random.default_rng(int(10 * nu))
    n, d, m = 400, 3, 15
    X = rng.random((n, d))
    y = np.sin(4 * X[:, 0]) + 0.3 * rng.standard_normal(n)
    theta = np.log([1.3, 0.3, 0.6, 1.1, 0.02])
    order, groups = _structure(X, theta, d, m)
    Z = O.design(X, trend)[order]
    Zarg = Z if Z.shape[1] else None
    kw = dict(var_penalty=1.0, log_var_target=0.1)
    ll, g, F, b = vecchia_profile_loglik(theta, X[order], y[order], Zarg, groups, nu, **kw)
    ll0, g0, F0, b0 = O.vecchia_profile_loglik(theta, X[order], y[order], Zarg, groups, nu, **kw)
    assert abs(ll - ll0) <= 1e-10 * abs(ll0)
    assert _rel(g, g0) < 1e-9 and _rel(F, F0) < 1e-9
    if b0.size:
        assert _rel(b, b0) < 1e-9


def test_estimated_nu_loglik_and_exact_gp(gpu):
    from pyscaledvecchia_amd.likelihood import vecchia_profile_loglik
    rng = np.random.default_rng(1)
    n, d, m = 300, 3, 12
    X = rng.random((n, d))
    y = rng.standard_normal(n)
    Z = np.ones((n, 1))
    theta = np.log([1.7, 0.35, 0.9, 0.5, 1.8, 0.05])
    order, groups = _structure(X, theta, d, m)
    args = (theta, X[order], y[order], Z[order], groups, None)
    ll, g, F, b = vecchia_profile_loglik(*args, estimate_nu=True)
    ll0, g0, F0, b0 = O.vecchia_profile_loglik(*args, estimate_nu=True)
    assert abs(ll - ll0) <= 1e-10 * abs(ll0)
    # the log-nu column is a central difference (step 1e-3) of two Bessel
    # evaluations on each side: rounding of K_nu enters divided by 2e-3
    assert _rel(g, g0) < 1e-8 and _rel(F, F0) < 1e-8
    # full conditioning: the exact GP (reference tests/test_likelihood.py:162-189)
    n2, d2 = 50, 2
    X2 = rng.random((n2, d2))
    y2 = rng.standard_normal(n2)
    th2 = np.log([1.2, 0.4, 0.3, 1.75, 0.02])
    o2, gr2 = _structure(X2, th2, d2, n2 - 1)
    ll2 = vecchia_profile_loglik(th2, X2[o2], y2[o2], None, gr2, None, estimate_nu=True)[0]
    var, lam, nuv, nug = math.exp(th2[0]), np.exp(th2[1:3]), math.exp(th2[3]), math.exp(th2[4])
    K = var * (O.matern_corr_general(cdist(X2 / lam, X2 / lam), nuv) + nug * np.eye(n2))
    L = np.linalg.cholesky(K)
    a = np.linalg.solve(L, y2)
    exact = -0.5 * (n2 * O.LOG2PI + 2 * np.log(np.diag(L)).sum() + a @ a)
    assert abs(ll2 - exact) < 1e-9 * abs(exact)


def _matern_draw(n=500, seed=0):
    """The reference's estimated-nu test data (tests/test_gp.py:149-168): a
    Matern(nu = 1.75) GP draw, ranges (0.25, 0.4), variance 1.5, nugget 0.02."""
    rng = np.random.default_rng(seed)
    X = rng.random((n, 2))
    lam = np.array([0.25, 0.4])
    K = 1.5 * (O.matern_corr_general(cdist(X / lam, X / lam), 1.75) + 0.02 * np.eye(n))
    y = np.linalg.cholesky(K + 1e-10 * np.eye(n)) @ rng.standard_normal(n)
    return X, y


@pytest.mark.parametrize("nu", [1.75, None])
def test_general_nu_fit_and_predict_match_oracle(gpu, nu):
    from pyscaledvecchia_amd import ScaledVecchiaGP
    X, y = _matern_draw()
    kw = dict(m_est=25, m_pred=50, n_est=500, nu=nu, trend="zero", random_state=0)
    g = ScaledVecchiaGP(**kw).fit(X, y)
    o = O.OracleGP(**kw).fit(X, y)
    print(f"\nnu={nu}: ll {g.loglik_:.8f} oracle {o.loglik_:.8f}; nu_ {g.nu_:.6f} {o.nu_:.6f}")
    assert g._estimate_nu == (nu is None)
    if nu is None:
        # reference tests/test_gp.py:170-174: a sane neighbourhood of the truth
        assert 0.5 < g.nu_ < 5.0 and g.relevance_[0] > g.relevance_[1]
    assert abs(g.loglik_ - o.loglik_) <= 1e-8 * abs(o.loglik_)
    np.testing.assert_allclose(g.theta_, o.theta_, rtol=1e-6, atol=1e-6)
    assert g.nu_ == pytest.approx(o.nu_, rel=1e-6)
    assert g.b_ == pytest.approx(o.b_, rel=1e-6)
    Xt = np.random.default_rng(21).random((40, 2))
    assert g.beta_.size == 0
    for a in ("theta_", "beta_", "b_", "_nu_fit"):
        setattr(o, a, getattr(g, a))
    mu, sd = g.predict(Xt, return_std=True)
    mu0, sd0 = o.predict(Xt, return_std=True)
    assert _rel(mu, mu0) < 1e-9 and _rel(sd, sd0) < 1e-8
    out = g.predict_joint(Xt[:10], exact_var=True)
    out0 = o.predict_joint(Xt[:10], exact_var=True)
    assert _rel(out["mean"], out0["mean"]) < 1e-8
    assert "estimated" in g.summary() if nu is None else "fixed" in g.summary()
