"""Pins for the NumPy oracle (CPU only).

Each test checks an oracle function against something other than its own
formula: a library routine, a dense closed form, finite differences or a
pure-Python brute force of the definition.  Seeds and shapes of the
reference's own tests (tests/test_likelihood.py, tests/test_ordering.py) are
reused where they exist.
"""

import math

import numpy as np
import pytest
import scipy.linalg
import scipy.special

from oracle import vecchia_np as O


# ---------------------------------------------------------------- Matern
@pytest.mark.parametrize("nu", [0.5, 1.5, 2.5])
def test_matern_closed_forms_match_bessel(nu):
    r = np.linspace(1e-6, 5, 200)
    c = math.sqrt(2 * nu) * r
    M_bessel = 2 ** (1 - nu) / scipy.special.gamma(nu) * c ** nu * scipy.special.kv(nu, c)
    np.testing.assert_allclose(O.matern_corr(r, nu), M_bessel, rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("nu", [0.5, 1.5, 2.5])
def test_matern_g_is_derivative_over_r(nu):
    r = np.linspace(0.05, 5, 100)
    h = 1e-6
    g_num = (O.matern_corr(r + h, nu) - O.matern_corr(r - h, nu)) / (2 * h) / r
    np.testing.assert_allclose(O.matern_g(r, nu), g_num, rtol=1e-7)


def test_matern_g_at_zero():
    assert O.matern_g(np.array([0.0]), 0.5)[0] == 0.0
    assert O.matern_g(np.array([0.0]), 1.5)[0] == -3.0
    assert O.matern_g(np.array([0.0]), 2.5)[0] == pytest.approx(-5.0 / 3.0)


# ---------------------------------------------------------------- ordering
def _maximin_bruteforce(X, first):
    """Definition, in pure Python: repeatedly take the point whose distance
    to the chosen set is largest (lowest index on ties)."""
    n, d = X.shape
    chosen = [first]
    remaining = set(range(n)) - {first}

    def dist(i, j):
        s = 0.0
        for l in range(d):
            t = X[i, l] - X[j, l]
            s = s + t * t
        return math.sqrt(s)

    while remaining:
        best, bj = -1.0, None
        for i in sorted(remaining):
            di = min(dist(i, j) for j in chosen)
            if di > best:
                best, bj = di, i
        chosen.append(bj)
        remaining.remove(bj)
    return np.array(chosen)


@pytest.mark.parametrize("seed,n,d", [(0, 40, 3), (5, 35, 2), (7, 30, 5)])
def test_maximin_matches_definition(seed, n, d):
    X = np.random.default_rng(seed).random((n, d))
    first = O.maximin_first(X)
    np.testing.assert_array_equal(O.maximin_order(X), _maximin_bruteforce(X, first))


def test_maximin_ties_take_lowest_index():
    # a regular grid has many exact distance ties
    g = np.arange(5.0)
    X = np.array([[a, b] for a in g for b in g])
    first = O.maximin_first(X)
    np.testing.assert_array_equal(O.maximin_order(X), _maximin_bruteforce(X, first))


def test_maximin_first_known_answers():
    """The first point (ordering.py:72) where the answer is fixed by geometry:
    equally spaced points on a line (odd n: the middle one; even n: the two
    middle ones tie and the lower index wins), a point placed at the centroid
    of random data, and the centroid shared by two rows (the lower index)."""
    for n in (7, 101):
        X = np.arange(n, dtype=float)[:, None] * np.array([[1.0, -2.0, 0.5]])
        assert O.maximin_first(X) == n // 2
    X = np.arange(10, dtype=float)[:, None] * np.array([[1.0, 3.0]])
    assert O.maximin_first(X) == 4
    rng = np.random.default_rng(8)
    X = rng.standard_normal((500, 4))
    X = X - X.mean(0)                  # centroid at 0 up to rounding
    X[123] = 0.0
    assert O.maximin_first(X) == 123
    X[40] = 0.0
    assert O.maximin_first(X) == 40


def test_maximin_permutation_and_deterministic():
    X = np.random.default_rng(0).random((200, 4))
    o = O.maximin_order(X)
    assert sorted(o.tolist()) == list(range(200))
    np.testing.assert_array_equal(o, O.maximin_order(X))


def test_find_ordered_nn_matches_brute_force():
    # reference tests/test_ordering.py:22-35 (rng(2), n=120, d=3, m=8)
    rng = np.random.default_rng(2)
    n, d, m = 120, 3, 8
    X = rng.random((n, d))
    Xo = X[O.maximin_order(X)]
    nn = O.find_ordered_nn(Xo, m)
    for i in range(1, n):
        k = min(i, m)
        dist = np.linalg.norm(Xo[:i] - Xo[i], axis=1)
        assert set(np.argsort(dist)[:k].tolist()) == set(nn[i, 1:k + 1].tolist())


def test_find_ordered_nn_structure_and_start():
    X = np.random.default_rng(3).random((30, 2))
    nn = O.find_ordered_nn(X, 5)
    assert np.array_equal(nn[:, 0], np.arange(30))
    assert np.all(nn[0, 1:] == -1)
    assert np.sum(nn[2, 1:] != -1) == 2
    X = np.random.default_rng(4).random((60, 2))
    full, part = O.find_ordered_nn(X, 6), O.find_ordered_nn(X, 6, start=40)
    np.testing.assert_array_equal(full[40:], part[40:])
    assert np.all(part[1:40, 1:] == -1)


def test_knn_query_matches_sorted_distances():
    rng = np.random.default_rng(9)
    Xr, Xq = rng.random((300, 3)), rng.random((20, 3))
    nb = O.knn_query(Xr, Xq, 12)
    for i in range(20):
        dist = np.linalg.norm(Xr - Xq[i], axis=1)
        np.testing.assert_array_equal(nb[i], np.argsort(dist, kind="stable")[:12])


def test_distance_rounding_is_sequential_without_fma():
    """The canonical squared distance rounds each product before the add: on
    near-tie pairs where either fused evaluation would pick the first
    candidate, the oracle picks the second (pins the no-FMA definition)."""
    from _helpers import near_tie_pairs
    Xr, Xq = near_tie_pairs()
    nb = O.knn_query(Xr, Xq, 1)
    np.testing.assert_array_equal(nb[:, 0], 2 * np.arange(len(Xq)) + 1)


# ---------------------------------------------------------------- blocks
@pytest.mark.parametrize("nu", [0.5, 1.5, 2.5])
def test_block_cov_derivatives_match_fd(nu):
    rng = np.random.default_rng(11)
    B, K, d = 3, 7, 3
    Xb = rng.random((B, K, d))
    theta = np.log([1.3, 0.4, 0.7, 2.0, 0.03])
    mask = (rng.random((B, K)) > 0.3).astype(float)

    def sig(th):
        return O.block_cov(Xb, np.exp(th[1:1 + d]), math.exp(th[0]), math.exp(th[-1]),
                           nu, derivs=False, nug_mask=mask)[0]

    S, dS = O.block_cov(Xb, np.exp(theta[1:1 + d]), math.exp(theta[0]), math.exp(theta[-1]),
                        nu, derivs=True, nug_mask=mask)
    np.testing.assert_allclose(S, sig(theta), rtol=0, atol=0)
    h = 1e-6
    for j in range(d + 2):
        tp, tm = theta.copy(), theta.copy()
        tp[j] += h
        tm[j] -= h
        num = (sig(tp) - sig(tm)) / (2 * h)
        np.testing.assert_allclose(dS[:, j], num, atol=1e-7)


def test_solves_match_scipy():
    rng = np.random.default_rng(12)
    A = rng.random((4, 9, 9))
    S = A @ A.transpose(0, 2, 1) + 9 * np.eye(9)
    L, p = O.batch_chol(S)
    assert p is None
    Bm = rng.random((4, 9, 3))
    X = O.forward_solve(L, Bm)
    Y = O.backsolve_LT(L, Bm)
    for b in range(4):
        np.testing.assert_allclose(L[b], np.linalg.cholesky(S[b]))
        np.testing.assert_allclose(X[b], scipy.linalg.solve_triangular(L[b], Bm[b], lower=True))
        np.testing.assert_allclose(Y[b], scipy.linalg.solve_triangular(L[b].T, Bm[b], lower=False))


def test_jitter_ladder_applies_per_chunk():
    # duplicate rows make a block singular: the whole chunk takes jitter
    rng = np.random.default_rng(13)
    Xb = rng.random((3, 5, 2))
    Xb[1, 3] = Xb[1, 2]
    S, _ = O.block_cov(Xb, np.array([0.5, 0.5]), 1.0, 0.0, 2.5, derivs=False)
    S[1] -= 1e-6 * np.eye(5)           # indefinite: needs jitter above 1e-6
    L, p = O.batch_chol(S)
    assert p == -5
    scale = np.einsum("bii->b", S) / 5
    np.testing.assert_allclose(L[0] @ L[0].T, S[0] + 10.0 ** p * scale[0] * np.eye(5))


# ---------------------------------------------------------------- likelihood
@pytest.fixture
def toy_problem():
    # reference tests/test_likelihood.py:31-39
    rng = np.random.default_rng(1)
    n, d, nu = 60, 3, 2.5
    X = rng.random((n, d))
    theta = np.array([math.log(1.7), math.log(0.35), math.log(0.9),
                      math.log(3.0), math.log(0.05)])
    y = rng.standard_normal(n)
    return X, y, theta, nu


def _structure(X, theta, m):
    d = X.shape[1]
    Xs = X / np.exp(theta[1:1 + d])
    order = O.maximin_order(Xs)
    nn = O.find_ordered_nn(Xs[order], m)
    return order, O.nn_groups(nn, m)


@pytest.mark.parametrize("with_trend", [False, True])
def test_vecchia_equals_exact_gp_at_full_conditioning(toy_problem, with_trend):
    X, y, theta, nu = toy_problem
    n = len(y)
    order, groups = _structure(X, theta, n - 1)
    Z = np.column_stack([np.ones(n), X]) if with_trend else None
    ll, g, F, beta = O.vecchia_profile_loglik(theta, X[order], y[order],
                                              None if Z is None else Z[order], groups, nu)
    ll_e, F_e, beta_e = O.exact_loglik(X, y, theta, nu, Z)
    assert abs(ll - ll_e) < 1e-9 * abs(ll_e)
    np.testing.assert_allclose(F, F_e, rtol=1e-7, atol=1e-9)
    if with_trend:
        np.testing.assert_allclose(beta, beta_e, rtol=1e-8, atol=1e-10)


@pytest.mark.parametrize("nu", [0.5, 1.5, 2.5])
def test_gradient_matches_finite_differences(toy_problem, nu):
    X, y, theta, _ = toy_problem
    n = len(y)
    order, groups = _structure(X, theta, 10)
    Z = np.column_stack([np.ones(n), X])

    def f(th):
        return O.vecchia_profile_loglik(th, X[order], y[order], Z[order], groups, nu,
                                        var_penalty=1.0, log_var_target=0.2)

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


def test_loglik_without_grad_matches(toy_problem):
    X, y, theta, nu = toy_problem
    order, groups = _structure(X, theta, 10)
    a = O.vecchia_profile_loglik(theta, X[order], y[order], None, groups, nu)
    b = O.vecchia_profile_loglik(theta, X[order], y[order], None, groups, nu, need_grad=False)
    assert b[1] is None and b[2] is None
    assert abs(a[0] - b[0]) < 1e-10


def test_fisher_per_block_matches_dense_difference(toy_problem):
    # one block's Fisher term = 1/2[tr(S^-1 S_j S^-1 S_k) - tr(Sc^-1 Sc_j Sc^-1 Sc_k)]
    X, y, theta, nu = toy_problem
    d = X.shape[1]
    idx = np.array([3, 17, 25, 40, 8])      # conditioning set then self
    Xb = X[idx]
    _, F_full, _ = O.exact_loglik(Xb, y[idx], theta, nu)
    _, F_c, _ = O.exact_loglik(Xb[:-1], y[idx][:-1], theta, nu)
    rows = np.array([4])
    groups = [(rows, np.array([[0, 1, 2, 3]]))]
    Xo = Xb
    _, _, F_v, _ = O.vecchia_profile_loglik(theta, Xo, y[idx], None, groups, nu)
    np.testing.assert_allclose(F_v, F_full - F_c, rtol=1e-8, atol=1e-10)


# ---------------------------------------------------------------- estimator
def _sine(n, d=2, seed=0):
    rng = np.random.default_rng(seed)
    X = rng.random((n, d))
    return X, np.sin(3 * X[:, 0]) + 0.1 * X[:, 1]


def test_oracle_fit_predict_recovers_function():
    X, y = _sine(300)
    gp = O.OracleGP(m_est=15, m_pred=30, n_est=300).fit(X, y)
    Xt, yt = _sine(200, seed=1)
    mu, sd = gp.predict(Xt, return_std=True)
    assert np.sqrt(np.mean((mu - yt) ** 2)) < 0.25
    assert np.all(sd >= 0)


def test_predict_scaled_matches_dense_conditional():
    X, y = _sine(150, seed=2)
    gp = O.OracleGP(m_est=10, m_pred=20, n_est=150, var_correction=False).fit(X, y)
    Xt = np.random.default_rng(3).random((10, 2))
    Xs = (Xt - gp._lo) / gp._span
    Z = O.design(gp.X_, gp.trend)
    resid = gp.y_ - Z @ gp.beta_
    mu, va = gp._predict_scaled(Xs, gp.X_, resid, 20)
    d = X.shape[1]
    th = gp.theta_
    lam, var, nug = np.exp(th[1:1 + d]), math.exp(th[0]), math.exp(th[-1])
    nb = O.knn_query(gp.X_ / lam, Xs / lam, 20)
    for i in range(10):
        P = np.vstack([gp.X_[nb[i]], Xs[i]])
        r = scipy.spatial.distance.cdist(P / lam, P / lam)
        C = var * O.matern_corr(r, 2.5)
        C[np.arange(20), np.arange(20)] += var * nug
        k = C[:20, 20]
        w = np.linalg.solve(C[:20, :20], k)
        assert mu[i] == pytest.approx(w @ resid[nb[i]], rel=1e-9, abs=1e-12)
        assert va[i] == pytest.approx(C[20, 20] - w @ k, rel=1e-7)


def test_predict_joint_exact_when_fully_conditioned():
    X, y = _sine(40, seed=8)
    gp = O.OracleGP(m_est=10, m_pred=20, n_est=40, var_correction=False,
                    nugget=1e-3).fit(X, y)
    # the joint factor conditions on min(m, n) points, so it is exact for a
    # single prediction point with m >= n
    Xt = np.random.default_rng(9).random((1, 2))
    out = gp.predict_joint(Xt, m=60, exact_var=True)
    Xs = (Xt - gp._lo) / gp._span
    d = 2
    th = gp.theta_
    lam, var, nug = np.exp(th[1:1 + d]), math.exp(th[0]), math.exp(th[-1])
    P = np.vstack([gp.X_, Xs])
    r = scipy.spatial.distance.cdist(P / lam, P / lam)
    C = var * O.matern_corr(r, 2.5)
    n = 40
    C[np.arange(n), np.arange(n)] += var * nug
    resid = gp.y_ - O.design(gp.X_, gp.trend) @ gp.beta_
    W = np.linalg.solve(C[:n, :n], C[:n, n:])
    mu = W.T @ resid + O.design(Xs, gp.trend) @ gp.beta_
    V = C[n:, n:] - C[n:, :n] @ W
    np.testing.assert_allclose(out["mean"], gp._ymu + gp._ysd * mu, rtol=1e-7, atol=1e-9)
    np.testing.assert_allclose(out["var"], gp._ysd ** 2 * np.diag(V), rtol=1e-5, atol=1e-10)


def test_fisher_scoring_reaches_stationary_point():
    X, y = _sine(120, seed=4)
    gp = O.OracleGP(m_est=10, m_pred=20, n_est=120, var_correction=False, tol=1e-8,
                    max_iter=60, nugget=1e-4).fit(X, y)
    st = gp.state_
    ll, g, M, _ = O.vecchia_profile_loglik(gp.theta_, st["Xo"], st["yo"], st["Zo"],
                                           st["groups"], 2.5, var_penalty=1.0,
                                           log_var_target=math.log(gp.y_.var()))
    g, M = g[:-1], M[:-1, :-1]            # nugget is held fixed
    assert abs(np.linalg.solve(M, g) @ g) < 1e-4
    assert ll == pytest.approx(gp.loglik_, rel=1e-12)
    assert gp.trace[-1][1] >= gp.trace[0][1]


import scipy.spatial  # noqa: E402  (used above)


# ---------------------------------------------------------------- b-step
def _dense_marginal(Xtr, r_tr, Xq, theta, m, nu=2.5):
    """Noise-free conditional mean/variance of each query given its m nearest
    training points (brute-force kNN in the scaled space, dense solve)."""
    d = Xtr.shape[1]
    lam, var, nug = np.exp(theta[1:1 + d]), math.exp(theta[0]), math.exp(theta[-1])
    mu, va = np.empty(len(Xq)), np.empty(len(Xq))
    for i, x in enumerate(Xq):
        dist = np.linalg.norm((Xtr - x) / lam, axis=1)
        nb = np.argsort(dist, kind="stable")[:m]
        P = np.vstack([Xtr[nb], x])
        C = var * O.matern_corr(scipy.spatial.distance.cdist(P / lam, P / lam), nu)
        C[np.arange(m), np.arange(m)] += var * nug
        w = np.linalg.solve(C[:m, :m], C[:m, m])
        mu[i] = w @ r_tr[nb]
        va[i] = C[m, m] - w @ C[:m, m]
    return mu, va


def test_estimate_b_split_and_minimiser():
    """b of the oracle fit equals the minimiser, found numerically, of the
    held-out negative log score sum_i log(b v_i) + (y_i - mu_i)^2 / (b v_i)
    (gp.py:410-425) on a split rebuilt by hand from the RNG call order:
    default_rng(random_state) -> choice(n, n_est) when n > n_est (gp.py:159-161)
    -> permutation(n), first min(2000, n // 5) held out (gp.py:413-415)."""
    import scipy.optimize
    X, y = _sine(300, seed=5)
    y = y + 0.05 * np.random.default_rng(6).standard_normal(300)
    gp = O.OracleGP(m_est=10, m_pred=25, n_est=200, trend="linear", random_state=3).fit(X, y)
    rng = np.random.default_rng(3)
    rng.choice(300, 200, replace=False)
    idx = rng.permutation(300)
    te, tr = idx[:60], idx[60:]
    Z = O.design(gp.X_, "linear")
    r_tr = gp.y_[tr] - Z[tr] @ gp.beta_
    mu, va = _dense_marginal(gp.X_[tr], r_tr, gp.X_[te], gp.theta_, 25)
    mu = mu + Z[te] @ gp.beta_
    e2 = (gp.y_[te] - mu) ** 2

    def score(lb):
        b = math.exp(lb)
        return float(np.sum(np.log(b * va) + e2 / (b * va)))
    res = scipy.optimize.minimize_scalar(score, bracket=(-3.0, 3.0), tol=1e-12)
    assert gp.b_ == pytest.approx(math.exp(res.x), rel=1e-6)
    # a different split (no subsample draw first) gives a different b
    rng2 = np.random.default_rng(3)
    te2 = rng2.permutation(300)[:60]
    assert not np.array_equal(np.sort(te2), np.sort(te))


def test_estimate_b_floor_and_small_n():
    X, y = _sine(150, seed=7)
    gp = O.OracleGP(m_est=8, m_pred=15, n_est=150).fit(X, y)
    assert gp.b_ == 1.0                       # n < 200: no correction (gp.py:237)
    X, y = _sine(250, seed=7)
    gp = O.OracleGP(m_est=8, m_pred=15, n_est=250, trend="linear").fit(X, y)
    gp.y_ = O.design(gp.X_, "linear") @ gp.beta_   # zero residuals -> b floored
    assert gp._estimate_b(np.random.default_rng(0)) == 1e-8


# ---------------------------------------------------------------- joint prediction
def _dense_vecchia_joint(gp, Xs, m, n_sim, seed):
    """The joint predictive law of the observed-first Vecchia approximation
    written from its definition: prediction point i (in the maximin order of
    the prediction points) regresses on its min(m, n) nearest predecessors
    among all observations and the earlier prediction points,
    y_i = b_i' y_N(i) + e_i, e_i ~ N(0, d_i), with observation neighbours noisy
    and prediction points noise-free (gp.py:321-339).  Then
    y_p = B_pp y_p + B_po y_o + e, so mean = (I - B_pp)^-1 B_po r_o and
    cov = (I - B_pp)^-1 D (I - B_pp)^-T; paths use the same standard normal
    draws as predict_joint (gp.py:389-397)."""
    n, d = gp.X_.shape
    ns = len(Xs)
    th = gp.theta_
    lam, var, nug = np.exp(th[1:1 + d]), math.exp(th[0]), math.exp(th[-1])
    Xp = Xs / lam
    first = int(np.argmin(((Xp - Xp.mean(0)) ** 2).sum(1)))
    porder = _maximin_bruteforce(Xp, first)
    Xall = np.vstack([gp.X_, Xs[porder]])      # observations, then ordered predictions
    noisy = np.r_[np.ones(n), np.zeros(ns)]
    k = min(m, n)
    B = np.zeros((ns, n + ns))
    D = np.empty(ns)
    for i in range(ns):
        row = n + i
        dist = np.linalg.norm((Xall[:row] - Xall[row]) / lam, axis=1)
        nb = np.argsort(dist, kind="stable")[:k]
        P = np.vstack([Xall[nb], Xall[row]])
        C = var * O.matern_corr(scipy.spatial.distance.cdist(P / lam, P / lam), 2.5)
        C[np.arange(k), np.arange(k)] += var * nug * noisy[nb]
        w = np.linalg.solve(C[:k, :k], C[:k, k])
        B[i, nb] = w
        D[i] = C[k, k] - w @ C[:k, k]
    resid = gp.y_ - O.design(gp.X_, gp.trend) @ gp.beta_
    IB = np.eye(ns) - B[:, n:]
    mu_ord = np.linalg.solve(IB, B[:, :n] @ resid)
    cov = np.linalg.solve(IB, np.linalg.solve(IB, np.diag(D)).T)
    inv = np.empty(ns, dtype=np.int64)
    inv[porder] = np.arange(ns)
    mu = mu_ord[inv] + O.design(Xs, gp.trend) @ gp.beta_
    eps = np.random.default_rng(seed).standard_normal((ns, n_sim))
    dr = np.linalg.solve(IB, np.sqrt(D)[:, None] * eps)[inv]
    samples = (mu[:, None] + math.sqrt(gp.b_) * dr).T
    return (gp._ymu + gp._ysd * mu, gp._ysd ** 2 * gp.b_ * np.diag(cov)[inv],
            gp._ymu + gp._ysd * samples)


@pytest.mark.parametrize("m", [6, 80])
def test_predict_joint_matches_dense_vecchia_definition(m):
    """Several prediction points: the prediction-point maximin order
    (gp.py:322), the neighbour sets among observations and earlier prediction
    points (gp.py:326, capped at min(m, n)), the noise-free mask of prediction
    neighbours (gp.py:339) and the three triangular solves (gp.py:380-397)
    against the dense definition.  m = 80 >= n + ns - 1 conditions every
    prediction point on n of its predecessors (the cap keeps it from all)."""
    X, y = _sine(60, seed=12)
    gp = O.OracleGP(m_est=10, m_pred=20, n_est=60, trend="constant", nugget=1e-3,
                    random_state=4).fit(X, y)
    Xt = np.random.default_rng(13).random((14, 2))
    out = gp.predict_joint(Xt, m=m, exact_var=True, n_sim=7, random_state=21)
    Xs = (Xt - gp._lo) / gp._span
    mean, var, samples = _dense_vecchia_joint(gp, Xs, m, 7, 21)
    np.testing.assert_allclose(out["mean"], mean, rtol=1e-9, atol=1e-11)
    np.testing.assert_allclose(out["var"], var, rtol=1e-8, atol=1e-12)
    np.testing.assert_allclose(out["samples"], samples, rtol=1e-9, atol=1e-10)
