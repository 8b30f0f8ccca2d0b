"""Learned warm start (the packaged network) on the GPU: what it gains over
the default start (gp.py:169-177), and the fallback to that start."""

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
    """A 'network' proposing absurd ranges: the start guard must reject it."""

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
    assert g.loglik_ == ref.loglik_
    np.testing.assert_array_equal(g.theta_, ref.theta_)
    # batched path: same guard
    b_ref = ScaledVecchiaGP(m_est=15, var_correction=False)
    b_bad = ScaledVecchiaGP(m_est=15, var_correction=False, warm_start=_BadStart())
    BatchedFit([b_ref, b_bad]).fit([X, X], [y, y])
    assert b_bad.warm_start_used_ is False
    assert b_bad.loglik_ == b_ref.loglik_
