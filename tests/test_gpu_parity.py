"""GPU parity: libsvh.so (through the C ABI) against the NumPy oracle.

Bars: ordering and neighbour sets bit-exact; float64 likelihood, gradient,
Fisher and beta within 1e-10 relative (1e-9 for Fisher/gradient entries,
relative to the largest entry); predictions within 1e-9 relative.
"""

import math

import numpy as np
import pytest

from oracle import vecchia_np as O

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


# ------------------------------------------------------------------ (a)
@pytest.mark.parametrize("n,d,seed", [(1, 3, 0), (2, 2, 1), (57, 3, 2), (300, 4, 3),
                                      (2000, 8, 4), (5000, 3, 5), (3000, 10, 6), (1500, 9, 7),
                                      (1200, 12, 8)])
def test_maximin_bit_exact(gpu, n, d, seed):
    X = np.random.default_rng(seed).random((n, d)) / 0.3
    first = O.maximin_first(X)
    np.testing.assert_array_equal(gpu.maximin(X, first), O.maximin_order(X, first))


def test_maximin_grid_ties_and_duplicates(gpu):
    g = np.arange(9.0)
    X = np.array([[a, b] for a in g for b in g])
    X = np.vstack([X, X[:7]])                      # exact duplicates too
    first = O.maximin_first(X)
    np.testing.assert_array_equal(gpu.maximin(X, first), O.maximin_order(X, first))


def test_maximin_large_prefix_and_property(gpu):
    n, d = 50000, 8
    X = np.random.default_rng(1002).random((n, d)) / (0.2 * math.sqrt(d))
    first = O.maximin_first(X)
    o = gpu.maximin(X, first)
    assert np.array_equal(np.sort(o), np.arange(n))
    ref = O.maximin_order(X, first, steps=1500)
    np.testing.assert_array_equal(o[:1500], ref)
    # the whole order against the C restatement of the oracle (~20 s on one core)
    from oracle import c_oracle as C
    if C.available():
        np.testing.assert_array_equal(o, C.maximin_order(X, first))
    # selected distances to the chosen set never increase
    nn = gpu.ordered_knn(X[o], 1)
    r = np.sqrt(((X[o][1:] - X[o][nn[1:, 1]]) ** 2).sum(1))
    assert np.all(np.diff(r) <= 1e-12)


def _outlier_slices(nslice, seed):
    """A tight cluster with four far outliers in every run of 256 consecutive
    rows: after the first pick every 256-row slice publishes its outliers
    above the bound, so a round has 4 * nslice survivors."""
    rng = np.random.default_rng(seed)
    X = rng.random((256 * nslice, 3)) * 0.1
    for s in range(nslice):
        rows = s * 256 + rng.choice(256, 4, replace=False)
        X[rows] = rng.random((4, 3)) * 10.0
    return X


def test_maximin_survivor_overflow_stays_exact(gpu, monkeypatch):
    """More survivors than the selection's capacity (forced small by
    SVH_MM_SORTCAP): those rounds accept only the largest candidate, and
    the order stays the serial one."""
    from oracle import c_oracle as C
    X = _outlier_slices(80, 21)
    first = O.maximin_first(X)
    ref = C.maximin_order(X, first) if C.available() else O.maximin_order(X, first)
    np.testing.assert_array_equal(gpu.maximin(X, first), ref)
    assert gpu.timing_read("maximin_fallbacks")[1] == 0
    monkeypatch.setenv("SVH_MM_SORTCAP", "256")
    np.testing.assert_array_equal(gpu.maximin(X, first), ref)
    assert gpu.timing_read("maximin_fallbacks")[1] > 0


@pytest.mark.parametrize("top", ["4", "8", "16"])
def test_maximin_both_list_depths_exact(gpu, monkeypatch, top):
    """Slices publish 4, 8 or 16 candidates (by default 16 up to 32 slices,
    8 up to 112); every depth gives the serial order, on random points and on
    a grid with exact ties."""
    from oracle import c_oracle as C
    monkeypatch.setenv("SVH_MM_TOP", top)
    X = np.random.default_rng(31).random((20000, 5)) / 0.3
    g = np.arange(12.0)
    G = np.array([[a, b, c] for a in g for b in g for c in g])
    for Y in (X, G):
        first = O.maximin_first(Y)
        ref = C.maximin_order(Y, first) if C.available() else O.maximin_order(Y, first)
        np.testing.assert_array_equal(gpu.maximin(Y, first), ref)


@pytest.mark.parametrize("n,d,m,start,seed", [(120, 3, 8, 0, 2), (2000, 4, 30, 0, 7),
                                              (600, 8, 30, 400, 8), (60, 2, 6, 40, 4),
                                              (300, 1, 5, 0, 9), (200, 5, 140, 0, 10),
                                              (800, 10, 30, 0, 11), (500, 9, 140, 100, 12)])
def test_ordered_knn_bit_exact(gpu, n, d, m, start, seed):
    X = np.random.default_rng(seed).random((n, d))
    np.testing.assert_array_equal(gpu.ordered_knn(X, m, start), O.find_ordered_nn(X, m, start))


@pytest.mark.parametrize("n,d,m,start", [(3000, 10, 140, 2500), (700, 14, 30, 650),
                                         (400, 3, 20, 0), (90, 2, 8, 90)])
def test_ordered_knn_rows_match_full_rows(gpu, n, d, m, start):
    """svh_ordered_knn_rows = rows start.. of svh_ordered_knn (tiled kernel for
    d <= 12, one-wave kernel beyond; start = n gives no rows)."""
    X = np.random.default_rng(n + d).random((n, d))
    full = gpu.ordered_knn(X, m, start)
    np.testing.assert_array_equal(gpu.ordered_knn_rows(X, m, start), full[start:])


def test_ordered_knn_ties(gpu):
    g = np.arange(12.0)
    X = np.array([[a, b] for a in g for b in g])
    np.testing.assert_array_equal(gpu.ordered_knn(X, 10), O.find_ordered_nn(X, 10))


@pytest.mark.parametrize("nref,nq,d,k", [(3000, 200, 4, 140), (500, 50, 10, 30), (141, 5, 3, 141)])
def test_knn_query_bit_exact(gpu, nref, nq, d, k):
    rng = np.random.default_rng(nref)
    Xr, Xq = rng.random((nref, d)), rng.random((nq, d))
    np.testing.assert_array_equal(gpu.knn_query(Xr, Xq, k), O.knn_query(Xr, Xq, k))


def test_knn_distance_is_not_fma_contracted(gpu):
    """Near-tie pairs whose nearest member flips if any squared-distance term
    is fused into an FMA: the kernels must round each product (as the oracle)."""
    from _helpers import near_tie_pairs
    Xr, Xq = near_tie_pairs()
    nb = gpu.knn_query(Xr, Xq, 1)
    np.testing.assert_array_equal(nb[:, 0], 2 * np.arange(len(Xq)) + 1)
    # ordered form and maximin share the distance routine: the same pairs as
    # one ordered set (each query last in its group of three)
    X = np.vstack([np.vstack([Xr[2 * k], Xr[2 * k + 1], Xq[k]]) for k in range(len(Xq))])
    np.testing.assert_array_equal(gpu.ordered_knn(X, 1), O.find_ordered_nn(X, 1))


# ------------------------------------------------------------------ (b)
def _toy(n=60, d=3, seed=1):
    rng = np.random.default_rng(seed)
    X = rng.random((n, d))
    theta = np.array([math.log(1.7)] + [math.log(v) for v in (0.35, 0.9, 3.0, 0.5, 1.2)[:d]]
                     + [math.log(0.05)])
    y = rng.standard_normal(n)
    return X, y, theta


def _structure(X, theta, m):
    d = X.shape[1]
    Xs = X / np.exp(theta[1:1 + d])
    order = O.maximin_order(Xs)
    nn = O.find_ordered_nn(Xs[order], m)
    return order, O.nn_groups(nn, m)


@pytest.mark.parametrize("nu", [0.5, 1.5, 2.5])
@pytest.mark.parametrize("trend", ["zero", "constant", "linear"])
@pytest.mark.parametrize("m", [10, 59])
def test_loglik_matches_oracle(gpu, nu, trend, m):
    from pyscaledvecchia_amd.likelihood import vecchia_profile_loglik
    X, y, theta = _toy()
    order, groups = _structure(X, theta, m)
    Z = O.design(X, trend)[order]
    Zarg = Z if Z.shape[1] else None
    args = (theta, X[order], y[order], Zarg, groups, nu)
    kw = dict(var_penalty=1.0, log_var_target=0.1)
    ll, g, F, b = vecchia_profile_loglik(*args, **kw)
    ll0, g0, F0, b0 = O.vecchia_profile_loglik(*args, **kw)
    assert abs(ll - ll0) <= 1e-10 * abs(ll0)
    assert _rel(g, g0) < 1e-9
    assert _rel(F, F0) < 1e-9
    if b0.size:
        assert _rel(b, b0) < 1e-9
    ll1, g1, F1, b1 = vecchia_profile_loglik(*args, need_grad=False, **kw)
    assert g1 is None and F1 is None
    assert abs(ll1 - ll0) <= 1e-10 * abs(ll0)


def test_loglik_exact_gp_at_full_conditioning(gpu):
    from pyscaledvecchia_amd.likelihood import vecchia_profile_loglik
    X, y, theta = _toy()
    order, groups = _structure(X, theta, 59)
    ll, _, F, _ = vecchia_profile_loglik(theta, X[order], y[order], None, groups, 2.5)
    ll_e, F_e, _ = O.exact_loglik(X, y, theta, 2.5)
    assert abs(ll - ll_e) < 1e-9 * abs(ll_e)
    np.testing.assert_allclose(F, F_e, rtol=1e-7, atol=1e-9)


def test_loglik_jitter_ladder_matches_oracle(gpu):
    """A block holding the same point twice with a zero nugget is exactly
    singular: the oracle (LAPACK) and the GPU both fail at level 0 (zero
    pivot) and succeed with jitter 1e-12 * tr/K on that chunk only."""
    from pyscaledvecchia_amd.likelihood import vecchia_profile_loglik
    rng = np.random.default_rng(21)
    n, d, m = 300, 2, 12
    X = rng.random((n, d))
    y = rng.standard_normal(n)
    theta = np.array([0.0, math.log(0.3), math.log(0.3), -np.inf])   # nugget exactly 0
    order, groups = _structure(X, theta, m)
    rows, nb = groups[-1]
    nb = nb.copy()
    nb[100, 1] = nb[100, 0]                     # one block repeats a neighbour
    groups = groups[:-1] + [(rows, nb)]
    args = (theta, X[order], y[order], np.ones((n, 1)), groups, 2.5)
    rep = {}
    ll0, g0, F0, b0 = O.vecchia_profile_loglik(*args, report=rep)
    assert [p for p in rep["jitter"] if p is not None] == [-12]
    ll, g, F, b = vecchia_profile_loglik(*args)
    assert gpu.last_jitter_power == -12
    assert abs(ll - ll0) <= 1e-6 * abs(ll0)
    assert _rel(g, g0) < 1e-4
    assert _rel(F, F0) < 1e-4


def test_loglik_levels_reset_after_jitter(gpu):
    """Chunk levels raised by an escalating evaluation do not leak into the
    next one on the same context; evaluations that follow a clean one (the
    clearing launch skipped) give the same results bit for bit."""
    from pyscaledvecchia_amd.likelihood import vecchia_profile_loglik
    rng = np.random.default_rng(22)
    n, d, m = 300, 2, 12
    X = rng.random((n, d))
    y = rng.standard_normal(n)
    th_sing = np.array([0.0, math.log(0.3), math.log(0.3), -np.inf])
    th_pos = np.array([0.0, math.log(0.3), math.log(0.3), math.log(1e-3)])
    order, groups = _structure(X, th_sing, m)
    rows, nb = groups[-1]
    nb = nb.copy()
    nb[100, 1] = nb[100, 0]
    groups = groups[:-1] + [(rows, nb)]
    data = (X[order], y[order], np.ones((n, 1)), groups, 2.5)
    r1 = vecchia_profile_loglik(th_pos, *data)
    assert gpu.last_jitter_power == 0
    vecchia_profile_loglik(th_sing, *data)
    assert gpu.last_jitter_power == -12
    r3 = vecchia_profile_loglik(th_pos, *data)
    assert gpu.last_jitter_power == 0
    r4 = vecchia_profile_loglik(th_pos, *data)
    for a, b in ((r1, r3), (r1, r4)):
        assert a[0] == b[0]
        for u, v in zip(a[1:], b[1:]):
            np.testing.assert_array_equal(u, v)


def test_loglik_c2_shape_against_oracle(gpu):
    """Full C2 dimensions (d=8, linear trend, m=30) on a 6000-row slice."""
    from pyscaledvecchia_amd import VecchiaStructure
    rng = np.random.default_rng(1002)
    n, d = 6000, 8
    X = rng.random((n, d))
    y = np.sin(2 * np.pi * X[:, 0] / 0.5) + 0.8 * np.sin(2 * np.pi * X[:, 1] / 0.7) \
        + 0.01 * rng.standard_normal(n)
    theta = np.concatenate([[0.0], np.full(d, math.log(0.2 * math.sqrt(d))), [math.log(0.1)]])
    order, groups = _structure(X, theta, 30)
    Z = O.design(X, "linear")[order]
    st = VecchiaStructure(X[order], y[order], Z, groups)
    ll, g, F, b = st(theta, 2.5, var_penalty=1.0, log_var_target=0.0)
    ll0, g0, F0, b0 = O.vecchia_profile_loglik(theta, X[order], y[order], Z, groups, 2.5,
                                               var_penalty=1.0, log_var_target=0.0)
    assert abs(ll - ll0) <= 1e-10 * abs(ll0)
    assert _rel(g, g0) < 1e-9 and _rel(F, F0) < 1e-9 and _rel(b, b0) < 1e-9


# ------------------------------------------------------------------ (c)
def test_predict_matches_oracle(gpu):
    rng = np.random.default_rng(31)
    n, d, ns, m = 2000, 4, 300, 140
    X = rng.random((n, d))
    y = np.sin(4 * X[:, 0]) + X[:, 1] * X[:, 2]
    gpo = O.OracleGP(m_est=10, m_pred=m, n_est=n, max_iter=1, var_correction=False).fit(X, y)
    resid = gpo.y_ - O.design(gpo.X_, gpo.trend) @ gpo.beta_
    Xt = rng.random((ns, d))
    mu0, va0 = gpo._predict_scaled(Xt, gpo.X_, resid, m)
    mu, va, nb = gpu.predict(gpo.X_, resid, Xt, m, gpo.theta_, 2.5, return_nb=True)
    np.testing.assert_array_equal(nb, O.knn_query(gpo.X_ / gpo.ranges_, Xt / gpo.ranges_, m))
    assert _rel(mu, mu0) < 1e-9
    assert _rel(va, va0) < 1e-8


@pytest.mark.parametrize("m", [1, 15, 16, 31, 140, 142, 150])
def test_predict_block_sizes_match_oracle(gpu, m):
    """Block sizes around the 16-row tiles of the matrix-core kernel: the
    appended residual row in the target's tile (m = 140, 142) or the next
    (m = 15, 31), the 144-row limit (m = 142) and the packed kernel beyond it
    (m = 150)."""
    rng = np.random.default_rng(40 + m)
    n, d, ns = 1200, 3, 64
    X = rng.random((n, d))
    y = np.sin(4 * X[:, 0]) + X[:, 1] * X[:, 2]
    gpo = O.OracleGP(m_est=10, m_pred=m, n_est=n, max_iter=1, var_correction=False).fit(X, y)
    resid = gpo.y_ - O.design(gpo.X_, gpo.trend) @ gpo.beta_
    Xt = rng.random((ns, d))
    mu0, va0 = gpo._predict_scaled(Xt, gpo.X_, resid, m)
    mu, va = gpu.predict(gpo.X_, resid, Xt, m, gpo.theta_, 2.5)
    assert _rel(mu, mu0) < 1e-9
    assert _rel(va, va0) < 1e-8
    # observed targets (nugget on the target too)
    mu0, va0 = gpo._predict_scaled(Xt, gpo.X_, resid, m, noise_free=False)
    mu, va = gpu.predict(gpo.X_, resid, Xt, m, gpo.theta_, 2.5, noise_free=False)
    assert _rel(mu, mu0) < 1e-9
    assert _rel(va, va0) < 1e-8


def test_predict_mixed_precision_error_is_small(gpu):
    """FP32 matrix-core trailing updates (mixed mode) at m = 140: the
    predictive mean and variance stay within 1e-4 relative of FP64."""
    rng = np.random.default_rng(33)
    n, d, ns, m = 3000, 4, 200, 140
    X = rng.random((n, d))
    resid = np.sin(4 * X[:, 0]) + X[:, 1] * X[:, 2]
    theta = np.log([1.0, 0.3, 0.4, 0.8, 2.0, 1e-3])
    Xt = rng.random((ns, d))
    mu, va = gpu.predict(X, resid, Xt, m, theta, 2.5)
    mu2, va2 = gpu.predict(X, resid, Xt, m, theta, 2.5, precision="mixed")
    assert _rel(mu2, mu) < 1e-4
    assert _rel(va2, va) < 1e-4


def test_predict_jitter_on_duplicate_training_points(gpu):
    """Training points repeated exactly with a zero nugget make the prediction
    blocks singular: the chunk jitter ladder (1e-12 ... ) applies as in the
    oracle."""
    rng = np.random.default_rng(34)
    n, d, ns, m = 400, 2, 30, 140
    X0 = rng.random((n // 2, d))
    X = np.vstack([X0, X0])
    resid = np.sin(3 * X[:, 0])
    theta = np.array([0.0, math.log(0.3), math.log(0.5), math.log(1e-300)])   # nugget ~ 0
    gpo = O.OracleGP(m_pred=m, nugget=None)
    gpo.theta_, gpo.d_ = theta, d
    gpo._nu_fit = gpo.nu = 2.5
    Xt = rng.random((ns, d))
    mu0, va0 = gpo._predict_scaled(Xt, X, resid, m)
    mu, va = gpu.predict(X, resid, Xt, m, theta, 2.5)
    assert _rel(mu, mu0) < 1e-5
    assert _rel(va, va0) < 1e-5


def test_block_columns_at_m140_match_oracle(gpu):
    rng = np.random.default_rng(35)
    n, d, ns, m = 600, 3, 24, 140
    X = rng.random((n, d))
    theta = np.log([1.3, 0.4, 0.6, 0.9, 1e-3])
    Xall = np.vstack([X, rng.random((ns, d))])
    Xsc = Xall / np.exp(theta[1:1 + d])
    order = np.concatenate([O.maximin_order(Xsc[:n]), O.maximin_order(Xsc[n:]) + n])
    nn = O.find_ordered_nn(Xsc[order], m, start=n)
    rows = np.arange(n, n + ns)
    blk = np.concatenate([nn[rows, 1:m + 1], rows[:, None]], axis=1)
    mask = np.ones((ns, m + 1))
    mask[:, m] = 0.0
    mask[:, :m] = (blk[:, :m] < n)
    a = gpu.block_columns(Xall[order], blk, mask, theta, 2.5)
    Xb = Xall[order][blk]
    Sig, _ = O.block_cov(Xb, np.exp(theta[1:1 + d]), math.exp(theta[0]), math.exp(theta[-1]),
                         2.5, derivs=False, nug_mask=mask)
    L, _ = O.batch_chol(Sig)
    E = np.zeros((ns, m + 1, 1))
    E[:, m, 0] = 1.0
    a0 = O.backsolve_LT(L, E)[:, :, 0]
    assert _rel(a, a0) < 1e-9


def test_block_columns_match_oracle(gpu):
    rng = np.random.default_rng(32)
    n, d, ns, m = 500, 3, 40, 30
    X = rng.random((n, d))
    y = np.cos(3 * X[:, 0]) + X[:, 2]
    gpo = O.OracleGP(m_est=10, m_pred=m, n_est=n, max_iter=1, var_correction=False).fit(X, y)
    Xs = rng.random((ns, d))
    U_op, U_pp, perm, oo = gpo._joint_factor(Xs, m)
    # same blocks through the GPU
    Xall = np.vstack([gpo.X_, Xs])
    Xsc = Xall / gpo.ranges_
    order = np.concatenate([O.maximin_order(Xsc[:n]), O.maximin_order(Xsc[n:]) + n])
    nn = O.find_ordered_nn(Xsc[order], m, start=n)
    rows = np.arange(n, n + ns)
    blk = np.concatenate([nn[rows, 1:m + 1], rows[:, None]], axis=1)
    mask = np.ones((ns, m + 1))
    mask[:, m] = 0.0
    mask[:, :m] = (blk[:, :m] < n)
    a = gpu.block_columns(Xall[order], blk, mask, gpo.theta_, 2.5)
    U = np.zeros((n + ns, ns))
    for j in range(ns):
        U[blk[j], j] = a[j]
    Uref = np.vstack([U_op.toarray(), U_pp.toarray()])
    assert _rel(U, Uref) < 1e-9


# ------------------------------------------------------------------ estimator
def _sine(n, d=2, seed=0):
    rng = np.random.default_rng(seed)
    X = rng.random((n, d))
    return X, np.sin(3 * X[:, 0]) + 0.1 * X[:, 1]


def test_fit_matches_oracle_fit(gpu):
    from pyscaledvecchia_amd import ScaledVecchiaGP
    X, y = _sine(400, d=3, seed=5)
    kw = dict(m_est=15, m_pred=30, n_est=300, nugget=1e-4, random_state=3)
    g = ScaledVecchiaGP(**kw).fit(X, y)
    o = O.OracleGP(**kw).fit(X, y)
    assert abs(g.loglik_ - o.loglik_) <= 1e-8 * abs(o.loglik_)
    np.testing.assert_allclose(g.theta_, o.theta_, rtol=1e-6, atol=1e-6)
    assert g.b_ == pytest.approx(o.b_, rel=1e-6)
    Xt = np.random.default_rng(6).random((50, 3))
    mu, sd = g.predict(Xt, return_std=True)
    mu0, sd0 = o.predict(Xt, return_std=True)
    np.testing.assert_allclose(mu, mu0, rtol=1e-6, atol=1e-8)
    np.testing.assert_allclose(sd, sd0, rtol=1e-6)


def test_reference_style_api(gpu):
    from pyscaledvecchia_amd import ScaledVecchiaGP
    with pytest.raises(ValueError):
        ScaledVecchiaGP().fit(np.random.rand(10, 2), np.random.rand(9))
    with pytest.raises(RuntimeError):
        ScaledVecchiaGP().predict(np.zeros((1, 2)))
    with pytest.raises(ValueError):
        ScaledVecchiaGP(trend="bogus").fit(np.random.rand(20, 2), np.random.rand(20))
    X, y = _sine(300, seed=0)
    gp = ScaledVecchiaGP(m_est=15, m_pred=30, nu=2.5, n_est=300).fit(X, y)
    Xt, yt = _sine(200, seed=1)
    mu, sd = gp.predict(Xt, return_std=True)
    assert np.sqrt(np.mean((mu - yt) ** 2)) < 0.25 and np.all(sd >= 0)
    mu2, var = gp.predict(Xt, return_var=True)
    np.testing.assert_allclose(mu, mu2)
    np.testing.assert_allclose(sd ** 2, var, rtol=1e-10)
    np.testing.assert_allclose(gp.predict(Xt), mu)
    assert "ScaledVecchiaGP" in gp.summary() and "loglik" in gp.summary()
    g2 = ScaledVecchiaGP(m_est=10, m_pred=20, n_est=120, nugget=1e-8).fit(*_sine(120, seed=6))
    assert g2.nugget_ == pytest.approx(1e-8, rel=1e-6)
    assert np.all(g2.ranges_ > 0) and np.all(g2.relevance_ > 0) and g2.variance_ > 0


@pytest.mark.parametrize("trend", ["zero", "constant", "linear"])
def test_all_trends(gpu, trend):
    from pyscaledvecchia_amd import ScaledVecchiaGP
    X, y = _sine(100, seed=13)
    gp = ScaledVecchiaGP(m_est=8, m_pred=15, n_est=100, trend=trend).fit(X, y)
    assert gp.predict(X[:5]).shape == (5,)


def test_joint_prediction(gpu):
    from pyscaledvecchia_amd import ScaledVecchiaGP
    X, y = _sine(200, seed=8)
    gp = ScaledVecchiaGP(m_est=15, m_pred=30, n_est=200).fit(X, y)
    Xt = np.random.default_rng(9).random((25, 2))
    joint = gp.predict_joint(Xt, m=30, exact_var=True)
    np.testing.assert_allclose(gp.predict(Xt), joint["mean"], atol=0.05)
    assert np.all(joint["var"] >= 0)
    s = gp.sample_joint(Xt[:12], n_sim=25, m=20)
    assert s.shape == (25, 12) and np.all(np.isfinite(s))


def test_device_rebuild_matches_host_structure(gpu):
    """svh_vecchia_rebuild (maximin + ordered kNN + structure on the device)
    gives the order and the objective of the host path (maximin_order,
    find_ordered_nn, _nn_groups, VecchiaStructure), bit for bit, also on a
    second rebuild that reuses the resident data."""
    from pyscaledvecchia_amd.likelihood import DeviceRebuild, VecchiaStructure
    from pyscaledvecchia_amd.ordering import _nn_groups, find_ordered_nn, maximin_order
    rng = np.random.default_rng(31)
    n, d, m = 3000, 5, 20
    Xe = rng.random((n, d))
    ye = rng.standard_normal(n)
    Ze = np.column_stack([np.ones(n), Xe])
    dev = DeviceRebuild(Xe, ye, Ze, m, ctx=gpu)
    theta = np.array([0.1, -0.5, -0.2, 0.3, 0.0, 1.0, math.log(0.05)])
    for shift, nu in ((0.0, 1.5), (0.3, 2.5), (-0.4, 0.5)):
        th = theta.copy()
        th[1:1 + d] += shift
        lam = np.exp(th[1:1 + d])
        Xs = Xe / lam
        order = maximin_order(Xs)
        nn = find_ordered_nn(Xs[order], m)
        host = VecchiaStructure(Xe[order], ye[order], Ze[order], _nn_groups(nn, m), ctx=gpu)
        ref = host(th, nu)
        got_order = dev(lam)
        np.testing.assert_array_equal(got_order, order)
        got = dev.structure(th, nu)
        assert got[0] == ref[0]
        for a, b in zip(got[1:], ref[1:]):
            np.testing.assert_array_equal(a, b)
        # value-only pass (other jitter chunking)
        assert dev.structure(th, nu, need_grad=False)[0] == host(th, nu, need_grad=False)[0]


@pytest.mark.parametrize("n,d", [(5000, 3), (4000, 7), (6000, 8), (3000, 10), (2500, 12),
                                 (1, 4), (50000, 8)])
def test_device_first_point_matches_numpy(gpu, n, d):
    """The rebuild's first point computed on the device (first = -1) is the
    index NumPy's argmin(((X - X.mean(0)) ** 2).sum(1)) gives for X = Xe / lam
    (ordering.py:72), including exact ties (duplicated rows) and ranges that
    spread the scales over decades."""
    from pyscaledvecchia_amd.likelihood import DeviceRebuild
    rng = np.random.default_rng(40 + d)
    Xe = rng.random((n, d))
    if n > 10:
        Xe[n // 3] = Xe[n // 5]          # a duplicated row
        c = Xe.mean(0)
        Xe[7] = c
        Xe[n - 2] = c                    # two rows at the same distance from the centroid
    ye = rng.standard_normal(n)
    m = min(10, n - 1)
    dev = DeviceRebuild(Xe, ye, np.ones((n, 1)), m, ctx=gpu)
    for lam in (np.full(d, 0.37), np.exp(rng.uniform(-3, 3, d))):
        order = dev(lam)
        assert order[0] == O.maximin_first(Xe / lam)   # the reference's one-pass expression


@pytest.mark.parametrize("n,d,m,trend", [(2500, 4, 15, "constant"), (400, 12, 10, "zero"),
                                         (60, 3, 59, "linear")])
def test_device_rebuild_fit_matches_reference_path(gpu, n, d, m, trend):
    """A whole fit with device rebuilds equals the fit through the host
    rebuild path (same theta, same log-likelihood, bit for bit): the C2-like
    case, d = 12 (widest maximin registers; the likelihood falls back from the
    rows kernel), and m = n - 1 (every earlier point a neighbour)."""
    from pyscaledvecchia_amd import ScaledVecchiaGP
    from pyscaledvecchia_amd import likelihood as LK
    rng = np.random.default_rng(5)
    X = rng.random((n, d))
    y = np.sin(5 * X[:, 0]) + X[:, 1] * X[:, 2] + 0.05 * rng.standard_normal(n)
    kw = dict(m_est=m, n_est=n, trend=trend, var_correction=False)
    g1 = ScaledVecchiaGP(**kw).fit(X, y)
    keep = LK.DeviceRebuild.supported
    try:
        LK.DeviceRebuild.supported = staticmethod(lambda d, m: False)
        g2 = ScaledVecchiaGP(**kw).fit(X, y)
    finally:
        LK.DeviceRebuild.supported = keep
    assert g1.loglik_ == g2.loglik_
    np.testing.assert_array_equal(g1.theta_, g2.theta_)
    assert g1.n_rebuilds_ == g2.n_rebuilds_


@pytest.mark.parametrize("n,ns,m,seed", [(300, 40, 30, 1), (2000, 1500, 40, 2), (500, 1, 20, 3)])
def test_joint_solve_matches_sparse_triangular(gpu, n, ns, m, seed):
    """svh_joint_solve (level-scheduled on the device) against
    scipy.sparse.linalg.spsolve_triangular -- the reference's own call,
    gp.py:380-401 -- on the same factor: mean, several right-hand sides and
    the exact variances (row sums of squares of Lt^-1; ns = 1500 spans two
    identity chunks)."""
    from scipy import sparse
    from scipy.sparse.linalg import spsolve_triangular
    from pyscaledvecchia_amd import ScaledVecchiaGP
    rng = np.random.default_rng(seed)
    X = rng.random((n, 3))
    y = np.sin(4 * X[:, 0]) + X[:, 1] * X[:, 2] + 0.01 * rng.standard_normal(n)
    gp = ScaledVecchiaGP(m_est=15, m_pred=m, var_correction=False).fit(X, y)
    Xs = rng.random((ns, 3))
    block, av, perm, order_obs = gp._joint_factor(Xs, m)
    nobs, K = gp.X_.shape[0], block.shape[1]
    resid = gp.y_ - gp._design(gp.X_) @ gp.beta_
    ro = resid[order_obs]
    U = sparse.coo_matrix((av.ravel(), (block.ravel(), np.repeat(np.arange(ns), K))),
                          shape=(nobs + ns, ns)).tocsc()
    Lt = U[nobs:, :].T.tocsr()
    ref_mu = spsolve_triangular(Lt, -(U[:nobs, :].T @ ro), lower=True)
    eps = rng.standard_normal((ns, 7))
    ref_x = spsolve_triangular(Lt, eps, lower=True)
    S = spsolve_triangular(Lt, np.eye(ns), lower=True)
    ref_var = (S ** 2).sum(1)
    mu, x, var = gpu.joint_solve(nobs, block, av, resid=ro, rhs=eps, exact_var=True)
    np.testing.assert_allclose(mu, ref_mu, rtol=1e-9, atol=1e-12 * np.abs(ref_mu).max())
    np.testing.assert_allclose(x, ref_x.reshape(ns, -1), rtol=1e-9,
                               atol=1e-12 * np.abs(ref_x).max())
    np.testing.assert_allclose(var, ref_var, rtol=1e-9)


def test_predict_joint_matches_oracle(gpu):
    """The whole joint prediction (mean, exact variance, sample paths with
    the same numpy RNG draws) against the oracle's predict_joint from the same
    fitted state.  The oracle orders the observations by maximin as
    gp.py:321 does; the estimator keeps their input order (the order changes
    no output)."""
    from pyscaledvecchia_amd import ScaledVecchiaGP
    X, y = _sine(400, seed=21)
    y = y + 0.05 * np.random.default_rng(23).standard_normal(len(y))
    gp = ScaledVecchiaGP(m_est=15, m_pred=30, n_est=400, trend="linear").fit(X, y)
    og = O.OracleGP(m_est=15, m_pred=30, n_est=400, trend="linear")
    for a in ("X_", "y_", "theta_", "beta_", "b_", "_lo", "_span", "_ymu", "_ysd", "d_"):
        setattr(og, a, getattr(gp, a))
    Xt = np.random.default_rng(22).random((60, 2))
    got = gp.predict_joint(Xt, m=30, exact_var=True, n_sim=9, random_state=5)
    ref = og.predict_joint(Xt, m=30, exact_var=True, n_sim=9, random_state=5)
    for k in ("mean", "samples"):
        np.testing.assert_allclose(got[k], ref[k], rtol=1e-8, atol=1e-10 * np.abs(ref[k]).max())
    # variances: relative to the prior variance of the fitted process (tiny
    # predictive variances are rounding-level differences of the factors)
    prior = gp._ysd ** 2 * math.exp(gp.theta_[0]) * gp.b_
    np.testing.assert_allclose(got["var"], ref["var"], rtol=1e-7, atol=1e-10 * prior)


def test_nu_property_matches_fixed_setting(gpu):
    """tests/test_gp.py:185-190 of the reference: nu is not estimated and nu_
    reports the fixed value."""
    from pyscaledvecchia_amd import ScaledVecchiaGP
    gp = ScaledVecchiaGP(m_est=8, m_pred=15, n_est=100, nu=2.5,
                         random_state=0).fit(*_sine(100, seed=23))
    assert gp._estimate_nu is False
    assert gp.nu_ == 2.5
