"""The benchmark configurations, each against the oracle on the GPU.

C1  n=2000, d=4 exact Matern-2.5 GP draw (lambda = 0.2, 0.4, 1, 20; nugget
    1e-4; mean 0.5), fit with m_est=30 and a constant trend, 500 predictions
    at m_pred=140: fit vs the oracle's fit, predictions vs the oracle's from
    the same fitted state.
C2  n=50,000, d=8 sin-sum function, linear trend, m=30: one full objective on
    the device-built structure vs the C restatement of the oracle, FP64 and
    mixed precision.
C3  prior draws of the batched configuration with d in 7..10, n >= 5000,
    nuggets 1e-8 .. 1e-2 and all three nu: batched evaluation vs the oracle,
    batched fits vs one-at-a-time fits.
C5  d=10, two-cluster inputs: 500 marginal predictions at m_pred=140 from
    200,000 training points vs the oracle; a 1,000-point joint path after
    20,000 observations (mean, exact variance, paths) vs the oracle.
Plus marginal prediction parity at nu = 0.5 and 1.5.
"""

import math
import os

import numpy as np
import pytest
from scipy.spatial.distance import cdist

from oracle import vecchia_np as O

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def c1_data():
    """rng(1001): X, X* uniform; y = 0.5 + chol(K) z, K = M_2.5(|(x - x')/lam|) + 1e-4 I."""
    rng = np.random.default_rng(1001)
    X = rng.random((2000, 4))
    Xt = rng.random((500, 4))
    lam = np.array([0.2, 0.4, 1.0, 20.0])
    K = O.matern_corr(cdist(X / lam, X / lam), 2.5) + 1e-4 * np.eye(2000)
    y = 0.5 + np.linalg.cholesky(K) @ rng.standard_normal(2000)
    return X, y, Xt


def _copy_state(src, dst):
    for a in ("X_", "y_", "theta_", "beta_", "b_", "_lo", "_span", "_ymu", "_ysd", "d_"):
        setattr(dst, a, getattr(src, a))


def test_c1_fit_and_predict_against_oracle(gpu):
    from pyscaledvecchia_amd import ScaledVecchiaGP
    X, y, Xt = c1_data()
    kw = dict(m_est=30, m_pred=140, n_est=2000, nu=2.5, trend="constant")
    g = ScaledVecchiaGP(**kw).fit(X, y)
    o = O.OracleGP(**kw).fit(X, y)
    assert abs(g.loglik_ - o.loglik_) <= 1e-8 * abs(o.loglik_)
    np.testing.assert_allclose(g.theta_, o.theta_, rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(g.beta_, o.beta_, rtol=1e-6, atol=1e-8)
    assert g.n_evals_ == o.n_evals and g.n_rebuilds_ == o.n_rebuilds
    assert g.b_ == pytest.approx(o.b_, rel=1e-6)
    # whole pipeline (each side from its own fit)
    mu, sd = g.predict(Xt, return_std=True)
    mu0, sd0 = o.predict(Xt, return_std=True)
    assert _rel(mu, mu0) < 1e-6 and _rel(sd, sd0) < 1e-6
    # prediction parity from one fitted state
    _copy_state(g, o)
    mu0, sd0 = o.predict(Xt, return_std=True)
    assert _rel(mu, mu0) < 1e-9
    assert _rel(sd, sd0) < 1e-8


def test_c2_full_objective_against_c_oracle(gpu):
    """One objective at the fit's starting point on the n = 50,000 structure
    the device rebuild produces (maximin order returned by the device, the C
    oracle's own neighbour lists), FP64 and mixed precision."""
    from bench import c2_data
    from oracle import c_oracle as C
    from pyscaledvecchia_amd.likelihood import DeviceRebuild
    if not C.available():
        pytest.fail("oracle/liboracle_c.so not built")
    X, y = c2_data()
    n, d, m = X.shape[0], 8, 30
    lo, span = X.min(0), X.max(0) - X.min(0)
    Xe = (X - lo) / span
    ye = (y - y.mean()) / y.std()
    Ze = np.column_stack([np.ones(n), Xe])
    theta = np.concatenate([[math.log(ye.var())], np.full(d, math.log(0.2 * math.sqrt(d))),
                            [math.log(0.1)]])
    dev = DeviceRebuild(Xe, ye, Ze, m, ctx=gpu)
    order = dev(np.exp(theta[1:1 + d]))
    lvt = math.log(ye.var())
    kw = dict(var_penalty=1.0, log_var_target=lvt)
    ll, g, F, b = dev.structure(theta, 2.5, **kw)
    Xs = Xe[order] / np.exp(theta[1:1 + d])
    nn = C.find_ordered_nn(Xs, m)
    groups = O.nn_groups(nn, m)
    nthr = min(16, os.cpu_count() or 1)
    ll0, g0, F0, b0 = C.vecchia_profile_loglik(theta, Xe[order], ye[order], Ze[order], groups,
                                               2.5, nthreads=nthr, **kw)
    assert abs(ll - ll0) <= 1e-10 * abs(ll0)
    assert _rel(g, g0) < 1e-9 and _rel(F, F0) < 1e-9 and _rel(b, b0) < 1e-9
    llm, gm, Fm, bm = dev.structure(theta, 2.5, precision="mixed", **kw)
    # mixed precision: FP32 derivative terms, W solve and products; the
    # log-likelihood and beta stay FP64
    assert llm == ll and np.array_equal(bm, b)
    assert _rel(gm, g0) < 1e-4 and _rel(Fm, F0) < 1e-4
    # and with the off-diagonal covariance entries assembled in FP32
    llc, gc, Fc, bc = dev.structure(theta, 2.5, precision="mixed_f32cov", **kw)
    assert abs(llc - ll0) <= 1e-6 * abs(ll0)
    assert _rel(gc, g0) < 1e-4 and _rel(Fc, F0) < 1e-4 and _rel(bc, b0) < 1e-5


def _c3_prior_draws(specs, seed=1003):
    """Datasets from the batched configuration's prior (X float32-valued,
    log10 lambda ~ U[-1, 1.5], sigma^2 = 1) with the given (n, d, nu, tau);
    y by 2048 random Fourier features of the Matern spectral density plus
    the nugget noise."""
    rng = np.random.default_rng(seed)
    out = []
    for n, d, nu, tau in specs:
        lam = 10.0 ** rng.uniform(-1, 1.5, d)
        X = rng.random((n, d)).astype(np.float32).astype(np.float64)
        F = 2048
        w = rng.standard_normal((F, d)) * np.sqrt(2 * nu / rng.chisquare(2 * nu, (F, 1)))
        bb = rng.uniform(0, 2 * math.pi, F)
        y = math.sqrt(2.0 / F) * np.cos((X / lam) @ w.T + bb).sum(1)
        y += math.sqrt(tau) * rng.standard_normal(n)
        out.append((X, y, nu))
    return out


C3_SPECS = [(5200, 10, 2.5, 1e-8), (6100, 7, 0.5, 1e-2), (5000, 9, 1.5, 1e-5)]


def test_c3_prior_batched_eval_against_oracle(gpu):
    from pyscaledvecchia_amd import ScaledVecchiaGP, _lib
    from pyscaledvecchia_amd.batch import BatchedFit
    data = _c3_prior_draws(C3_SPECS)
    m = 30
    models = [ScaledVecchiaGP(m_est=m, nu=nu, trend="constant", n_est=20000) for _, _, nu in data]
    bf = BatchedFit(models)
    preps = [g._prepare(X, y) for g, (X, y, _) in zip(models, data)]
    structs = [bf._structure(g, p, p["theta"], m) for g, p in zip(models, preps)]
    bh = _lib.BatchHandle(gpu)
    nd = len(data)
    bh.setup([p["d"] for p in preps], [1] * nd, [{0.5: 0, 1.5: 1, 2.5: 2}[nu] for *_, nu in data],
             [len(s["yo"]) for s in structs], np.concatenate([s["Xo"].ravel() for s in structs]),
             np.concatenate([s["yo"] for s in structs]),
             np.concatenate([s["Zo"].ravel() for s in structs]),
             np.concatenate([s["nbr"] for s in structs]),
             np.concatenate([s["kcnt"] for s in structs]),
             np.concatenate([s["chunk"] for s in structs]), [s["nchunk"] for s in structs], m,
             [p["lvt"] for p in preps], [0] * nd, [0.0] * nd, 1e3)
    th = np.zeros((nd, 12))
    for i, (p, spec) in enumerate(zip(preps, C3_SPECS)):
        th[i, :p["d"] + 2] = p["theta"]
        th[i, p["d"] + 1] = math.log(max(spec[3], 1e-8))     # nugget at the drawn value
    ll, g, F, beta, jr = bh.eval(th)
    for i, (s, p, (_, _, nu)) in enumerate(zip(structs, preps, data)):
        P = p["d"] + 2
        groups = O.nn_groups(np.column_stack([np.arange(len(s["yo"])), s["nbr"][:, :m]]), m)
        ll0, g0, F0, b0 = O.vecchia_profile_loglik(th[i, :P], s["Xo"], s["yo"], s["Zo"], groups,
                                                   nu, var_penalty=1.0, log_var_target=p["lvt"])
        assert abs(ll[i] - ll0) <= 1e-10 * abs(ll0), (i, ll[i], ll0)
        assert _rel(g[i, :P], g0) < 1e-8
        assert _rel(F[i, :P * P].reshape(P, P), F0) < 1e-8
        assert _rel(beta[i, :1], b0) < 1e-8


def test_c3_prior_batched_fit_equals_single_fits(gpu):
    from pyscaledvecchia_amd import ScaledVecchiaGP
    from pyscaledvecchia_amd.batch import BatchedFit
    data = _c3_prior_draws(C3_SPECS, seed=1004)
    kw = dict(m_est=30, n_est=5000, trend="constant", var_correction=False)
    singles = [ScaledVecchiaGP(nu=nu, **kw).fit(X, y) for X, y, nu in data]
    batch = [ScaledVecchiaGP(nu=nu, **kw) for _, _, nu in data]
    BatchedFit(batch, use_graph=True).fit([X for X, _, _ in data], [y for _, y, _ in data])
    for s, b in zip(singles, batch):
        assert b.loglik_ == pytest.approx(s.loglik_, rel=1e-9)
        np.testing.assert_allclose(b.theta_, s.theta_, rtol=1e-7, atol=1e-7)
        assert b.n_evals_ == s.n_evals_


def _fitted_state(gp, X, y, theta, beta=0.0):
    gp._lo, gp._span = X.min(0), X.max(0) - X.min(0)
    gp._ymu, gp._ysd = y.mean(), y.std()
    gp.X_, gp.y_ = (X - gp._lo) / gp._span, (y - gp._ymu) / gp._ysd
    gp.d_ = X.shape[1]
    gp.theta_ = np.asarray(theta, float)
    gp.beta_ = np.array([beta])
    gp.b_ = 1.3
    gp._nu_fit = gp.nu


def test_c5_marginal_prediction_against_oracle(gpu):
    from bench import c5_inputs
    from pyscaledvecchia_amd import ScaledVecchiaGP
    d, m = 10, 140
    X, y, Xs = c5_inputs(200_000, 500, d)
    theta = np.concatenate([[0.1], np.log([0.3, 0.4, 0.5, 0.8, 1.0, 1.5, 2.0, 3.0, 5.0, 8.0]),
                            [math.log(1e-4)]])
    gp = ScaledVecchiaGP(m_pred=m, nu=2.5)
    _fitted_state(gp, X, y, theta, beta=0.02)
    og = O.OracleGP(m_pred=m, nu=2.5)
    _copy_state(gp, og)
    mu, sd = gp.predict(Xs, return_std=True)
    mu0, sd0 = og.predict(Xs, return_std=True)
    assert _rel(mu, mu0) < 1e-9
    assert _rel(sd, sd0) < 1e-8


def test_c5_joint_path_against_oracle(gpu):
    from bench import c5_inputs
    from pyscaledvecchia_amd import ScaledVecchiaGP
    d, m = 10, 140
    X, y, Xs = c5_inputs(20_000, 1000, d, seed=1006)
    theta = np.concatenate([[0.0], np.full(d, math.log(0.5)), [math.log(1e-4)]])
    gp = ScaledVecchiaGP(m_pred=m, nu=2.5, random_state=0)
    _fitted_state(gp, X, y, theta)
    og = O.OracleGP(m_pred=m, nu=2.5, random_state=0)
    _copy_state(gp, og)
    got = gp.predict_joint(Xs, exact_var=True, n_sim=20, random_state=7)
    ref = og.predict_joint(Xs, exact_var=True, n_sim=20, random_state=7)
    for k in ("mean", "samples"):
        np.testing.assert_allclose(got[k], ref[k], rtol=1e-8, atol=1e-10 * np.abs(ref[k]).max())
    prior = gp._ysd ** 2 * math.exp(gp.theta_[0]) * gp.b_
    np.testing.assert_allclose(got["var"], ref["var"], rtol=1e-7, atol=1e-10 * prior)


@pytest.mark.parametrize("nu", [0.5, 1.5])
def test_predict_other_smoothness_against_oracle(gpu, nu):
    rng = np.random.default_rng(int(nu * 10))
    n, d, ns = 3000, 4, 200
    X = rng.random((n, d))
    resid = np.sin(4 * X[:, 0]) + X[:, 1] * X[:, 2]
    theta = np.log([1.2, 0.3, 0.4, 0.8, 2.0, 1e-3])
    Xt = rng.random((ns, d))
    og = O.OracleGP(m_pred=140, nu=nu)
    og.theta_, og.d_ = theta, d
    for m in (30, 140):
        mu0, va0 = og._predict_scaled(Xt, X, resid, m)
        mu, va = gpu.predict(X, resid, Xt, m, theta, nu)
        assert _rel(mu, mu0) < 1e-9
        assert _rel(va, va0) < 1e-8


@pytest.mark.parametrize("log_var", [math.log(1e6), math.log(1e-6)])
def test_predict_extreme_scales_and_tiny_nugget(gpu, log_var):
    """Prediction blocks with nugget 1e-12 and variance scales 1e6 / 1e-6: the
    pivots of the matrix-core kernel's diagonal tiles (hardware reciprocal
    square root refined by Newton steps) against the oracle."""
    rng = np.random.default_rng(91)
    n, d, ns, m = 2500, 3, 120, 140
    X = rng.random((n, d))
    resid = np.sin(4 * X[:, 0]) + X[:, 1] * X[:, 2]
    theta = np.array([log_var, math.log(0.5), math.log(0.7), math.log(1.1), math.log(1e-12)])
    Xt = rng.random((ns, d))
    og = O.OracleGP(m_pred=m, nu=2.5)
    og.theta_, og.d_ = theta, d
    mu0, va0 = og._predict_scaled(Xt, X, resid, m)
    mu, va = gpu.predict(X, resid, Xt, m, theta, 2.5)
    assert _rel(mu, mu0) < 1e-9
    assert _rel(va, va0) < 1e-8
