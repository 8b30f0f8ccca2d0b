"""Learned warm start (the packaged network) on the GPU: what it gains over
the default start (gp.py:169-177), and the guarantee that a learned start
never ends below the fit from that start (both fits run, the higher final
log-likelihood is kept)."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_learned_start_is_closer_to_the_fitted_parameters(gpu):
    """On held-out prior draws fitted from the default start, the network's
    log-ranges and log-variance are much closer to the fitted values than the
    default start log(0.2 sqrt d), log var(y) is; the starting point (ranges
    halfway between the two) is closer too."""
    from pyscaledvecchia_amd import ScaledVecchiaGP
    from pyscaledvecchia_amd import warmstart as W
    from pyscaledvecchia_amd.batch import BatchedFit
    ws = W.WarmStart.load()
    data = W.simulate_prior(gpu, 64, 5000, seed=123)
    models = [ScaledVecchiaGP(nu=nu, m_est=30, var_correction=False) for _, _, nu, _ in data]
    BatchedFit(models).fit([X for X, _, _, _ in data], [y for _, y, _, _ in data])
    e_net, e_ws, e_def, s_ws, s_def = [], [], [], [], []
    for g, (X, y, nu, _) in zip(models, data):
        prep = ScaledVecchiaGP(nu=nu, m_est=30)._prepare(X, y)
        d = X.shape[1]
        raw = ws.predict_raw(W.summary_features(prep["Xe"], prep["ye"], nu))[0]
        th_ws = ws.theta0(prep["Xe"], prep["ye"], nu)
        th_def = prep["theta_default"]
        e_net += list(np.clip(raw[1:1 + d], W.LOG_LAM_LO, W.LOG_LAM_HI) - g.theta_[1:1 + d])
        e_ws += list(th_ws[1:1 + d] - g.theta_[1:1 + d])
        e_def += list(th_def[1:1 + d] - g.theta_[1:1 + d])
        s_ws.append(th_ws[0] - g.theta_[0])
        s_def.append(th_def[0] - g.theta_[0])
    rms = lambda v: float(np.sqrt(np.mean(np.square(v))))   # noqa: E731
    assert rms(e_net) < 0.7 * rms(e_def), (rms(e_net), rms(e_def))
    assert rms(e_ws) < rms(e_def), (rms(e_ws), rms(e_def))
    assert rms(s_ws) < rms(s_def), (rms(s_ws), rms(s_def))


def test_warm_start_single_fit(gpu):
    from pyscaledvecchia_amd import ScaledVecchiaGP
    rng = np.random.default_rng(8)
    X = rng.random((3000, 4))
    y = np.sin(6 * X[:, 0]) + 0.5 * np.cos(3 * X[:, 1]) + 0.01 * rng.standard_normal(3000)
    g = ScaledVecchiaGP(m_est=20, m_pred=40, warm_start="learned").fit(X, y)
    assert isinstance(g.warm_start_used_, bool)
    assert np.isfinite(g.loglik_) and np.all(np.isfinite(g.theta_))
    mu = g.predict(X[:100])
    assert np.sqrt(np.mean((mu - y[:100]) ** 2)) < 0.1


class _BadStart:
    """A 'network' proposing absurd ranges: its fit must lose to the default."""

    def theta0(self, Xe, ye, nu):
        d = Xe.shape[1]
        return np.concatenate([[5.0], np.full(d, np.log(1e-6)), [np.log(0.1)]])

    def theta0_many(self, items):
        return [self.theta0(*t) for t in items]


def test_rejected_learned_start_reproduces_default_fit(gpu):
    from pyscaledvecchia_amd import ScaledVecchiaGP
    from pyscaledvecchia_amd.batch import BatchedFit
    rng = np.random.default_rng(9)
    X = rng.random((2000, 3))
    y = np.sin(5 * X[:, 0]) + X[:, 1] ** 2 + 0.01 * rng.standard_normal(2000)
    ref = ScaledVecchiaGP(m_est=15, var_correction=False).fit(X, y)
    g = ScaledVecchiaGP(m_est=15, var_correction=False, warm_start=_BadStart()).fit(X, y)
    assert g.warm_start_used_ is False
    # the kept fit is the batched fitter's default-start fit
    assert g.loglik_ == pytest.approx(ref.loglik_, rel=1e-9)
    np.testing.assert_allclose(g.theta_, ref.theta_, rtol=1e-7, atol=1e-7)
    # batched path: same guard
    b_ref = ScaledVecchiaGP(m_est=15, var_correction=False)
    b_bad = ScaledVecchiaGP(m_est=15, var_correction=False, warm_start=_BadStart())
    BatchedFit([b_ref, b_bad]).fit([X, X], [y, y])
    assert b_bad.warm_start_used_ is False
    assert b_bad.loglik_ == b_ref.loglik_


def test_learned_start_never_ends_below_default_c2(gpu):
    """C2 data (n = 50,000, d = 8, linear trend, m_est = 30): the learned-start
    fit's final log-likelihood is never below the default-start fit's."""
    import sys, os
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    from bench import c2_data
    from pyscaledvecchia_amd import ScaledVecchiaGP
    X, y = c2_data()
    kw = dict(m_est=30, m_pred=140, nu=2.5, trend="linear", n_est=50_000, var_correction=False)
    ref = ScaledVecchiaGP(**kw).fit(X, y)
    g = ScaledVecchiaGP(warm_start="learned", **kw).fit(X, y)
    print(f"\nC2: default {ref.loglik_:.6f} learned-start fit {g.loglik_:.6f} "
          f"(learned kept: {g.warm_start_used_})")
    assert g.loglik_ >= ref.loglik_ - 1e-9 * abs(ref.loglik_)
