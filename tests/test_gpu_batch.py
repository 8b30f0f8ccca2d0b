"""Batched fits (many uneven datasets in one launch) against one-at-a-time fits
and the oracle.  GPU only."""

import math

import numpy as np
import pytest

from oracle import vecchia_np as O

pytestmark = pytest.mark.gpu


def _datasets(seed, nd=6):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(nd):
        n = int(rng.integers(300, 1500))
        d = int(rng.integers(2, 7))
        X = rng.random((n, d))
        w = rng.standard_normal(d)
        y = np.sin(3 * X @ w) + 0.3 * X[:, 0] ** 2 + 0.05 * rng.standard_normal(n)
        out.append((X, y))
    return out


def test_batch_eval_matches_single_structure(gpu):
    from pyscaledvecchia_amd import _lib
    from pyscaledvecchia_amd.batch import BatchedFit
    from pyscaledvecchia_amd.gp import ScaledVecchiaGP
    data = _datasets(1, nd=5)
    nus = [0.5, 1.5, 2.5, 2.5, 1.5]
    trends = ["constant", "linear", "zero", "linear", "constant"]
    models = [ScaledVecchiaGP(m_est=20, nu=nu, trend=t, n_est=2000) for nu, t in zip(nus, trends)]
    bf = BatchedFit(models)
    preps = [g._prepare(X, y) for g, (X, y) in zip(models, data)]
    structs = [bf._structure(g, p, p["theta"], 20) for g, p in zip(models, preps)]
    bh = _lib.BatchHandle(gpu)
    q = [s["Zo"].shape[1] for s in structs]
    bh.setup([p["d"] for p in preps], q, [{0.5: 0, 1.5: 1, 2.5: 2}[nu] for nu in nus],
             [len(s["yo"]) for s in structs], np.concatenate([s["Xo"].ravel() for s in structs]),
             np.concatenate([s["yo"] for s in structs]),
             np.concatenate([s["Zo"].ravel() for s in structs]),
             np.concatenate([s["nbr"] for s in structs]),
             np.concatenate([s["kcnt"] for s in structs]),
             np.concatenate([s["chunk"] for s in structs]), [s["nchunk"] for s in structs], 20,
             [p["lvt"] for p in preps], [0] * 5, [0.0] * 5, 1e3)
    rng = np.random.default_rng(3)
    th = np.zeros((5, 12))
    for i, p in enumerate(preps):
        th[i, :p["d"] + 2] = p["theta"] + 0.3 * rng.standard_normal(p["d"] + 2)
    ll, g, F, beta, jr = bh.eval(th)
    for i, (s, p) in enumerate(zip(structs, preps)):
        P = p["d"] + 2
        groups = O.nn_groups(np.column_stack([np.arange(len(s["yo"])), s["nbr"][:, :20]]), 20)
        Zarg = s["Zo"] if q[i] else None
        ll0, g0, F0, b0 = O.vecchia_profile_loglik(th[i, :P], s["Xo"], s["yo"], Zarg, groups, nus[i])
        assert abs(ll[i] - ll0) <= 1e-10 * abs(ll0)
        np.testing.assert_allclose(g[i, :P], g0, rtol=1e-8, atol=1e-9 * np.abs(g0).max())
        np.testing.assert_allclose(F[i, :P * P].reshape(P, P), F0, rtol=1e-8,
                                   atol=1e-9 * np.abs(F0).max())
        if q[i]:
            np.testing.assert_allclose(beta[i, :q[i]], b0, rtol=1e-8, atol=1e-10)


@pytest.mark.parametrize("use_graph", [True, False])
def test_batched_fit_equals_individual_fits(gpu, use_graph):
    from pyscaledvecchia_amd import ScaledVecchiaGP
    from pyscaledvecchia_amd.batch import BatchedFit
    data = _datasets(7, nd=6)
    nus = [2.5, 1.5, 0.5, 2.5, 1.5, 2.5]
    kw = dict(m_est=15, m_pred=30, n_est=1000, trend="constant", var_correction=True)
    singles = [ScaledVecchiaGP(nu=nu, **kw).fit(X, y) for nu, (X, y) in zip(nus, data)]
    batch = [ScaledVecchiaGP(nu=nu, **kw) for nu in nus]
    bf = BatchedFit(batch, use_graph=use_graph)
    bf.fit([X for X, _ in data], [y for _, y in data])
    for s, b in zip(singles, batch):
        assert b.loglik_ == pytest.approx(s.loglik_, rel=1e-9)
        np.testing.assert_allclose(b.theta_, s.theta_, rtol=1e-7, atol=1e-7)
        np.testing.assert_allclose(b.beta_, s.beta_, rtol=1e-7, atol=1e-9)
        assert b.n_evals_ == s.n_evals_
        assert b.b_ == pytest.approx(s.b_, rel=1e-7)
    assert bf.stats["steps"] > 0


def test_batched_fit_fixed_nugget_and_linear_trend(gpu):
    from pyscaledvecchia_amd import ScaledVecchiaGP
    from pyscaledvecchia_amd.batch import fit_batch
    data = _datasets(11, nd=3)
    kw = dict(m_est=10, n_est=800, trend="linear", nugget=1e-6)
    models, stats = fit_batch([X for X, _ in data], [y for _, y in data], **kw)
    for m, (X, y) in zip(models, data):
        s = ScaledVecchiaGP(**kw).fit(X, y)
        assert m.nugget_ == pytest.approx(1e-6, rel=1e-9)
        assert m.loglik_ == pytest.approx(s.loglik_, rel=1e-9)
        np.testing.assert_allclose(m.predict(X[:10]), s.predict(X[:10]), rtol=1e-7, atol=1e-9)


def test_ordering_batch_bit_exact(gpu):
    rng = np.random.default_rng(5)
    Xs = [rng.random((n, d)) / 0.4 for n, d in [(700, 3), (1500, 8), (50, 2), (2100, 5)]]
    res = gpu.ordering_batch(Xs, 12)
    for X, (order, nn) in zip(Xs, res):
        ref = O.maximin_order(X)
        np.testing.assert_array_equal(order, ref)
        np.testing.assert_array_equal(nn, O.find_ordered_nn(X[ref], 12))


def _batch_run(bh, theta0, var_penalty, use_graph):
    """Fisher scoring to the end on a fixed structure (reorders declined)."""
    bh.fit_start(theta0, 40, 1e-4, var_penalty, use_graph)
    for _ in range(10000):
        st = bh.fit_step(1)
        for i in np.flatnonzero(st == 2):
            bh.set_reorder(int(i), False)
        if np.all((st == 3) | (st == 4)):
            break
    return bh.get()


def test_graph_recaptured_when_var_penalty_changes(gpu):
    """fit_start twice on one set-up with different var_penalty: the replayed
    graph must use the new value (same result as the uncaptured steps)."""
    from pyscaledvecchia_amd import _lib
    from pyscaledvecchia_amd.batch import BatchedFit
    from pyscaledvecchia_amd.gp import ScaledVecchiaGP
    data = _datasets(21, nd=2)
    models = [ScaledVecchiaGP(m_est=15, nu=2.5, trend="constant", n_est=2000) for _ in data]
    bf = BatchedFit(models)
    preps = [g._prepare(X, y) for g, (X, y) in zip(models, data)]
    structs = [bf._structure(g, p, p["theta"], 15) for g, p in zip(models, preps)]
    bh = _lib.BatchHandle(gpu)
    bh.setup([p["d"] for p in preps], [1, 1], [2, 2], [len(s["yo"]) for s in structs],
             np.concatenate([s["Xo"].ravel() for s in structs]),
             np.concatenate([s["yo"] for s in structs]),
             np.concatenate([s["Zo"].ravel() for s in structs]),
             np.concatenate([s["nbr"] for s in structs]),
             np.concatenate([s["kcnt"] for s in structs]),
             np.concatenate([s["chunk"] for s in structs]), [s["nchunk"] for s in structs], 15,
             [p["lvt"] - 1.0 for p in preps], [0, 0], [0.0, 0.0], 1e3)
    th0 = np.zeros((2, 12))
    for i, p in enumerate(preps):
        th0[i, :p["d"] + 2] = p["theta"]
    a = _batch_run(bh, th0, 1.0, True)       # captures the graph with var_penalty 1
    b = _batch_run(bh, th0, 10.0, True)      # must capture again
    c = _batch_run(bh, th0, 10.0, False)     # plain launches, var_penalty 10
    np.testing.assert_array_equal(b[0], c[0])
    np.testing.assert_array_equal(b[1], c[1])
    assert not np.array_equal(a[1], b[1])


def test_batch_setup_rejects_null_trend_with_columns(gpu):
    """The C ABI refuses Z = NULL with data given and q > 0 (no host read of a
    NULL pointer)."""
    import ctypes
    L = gpu._L
    i32 = lambda v: np.ascontiguousarray(v, np.int32)    # noqa: E731
    d, q, nu, nc = i32([2]), i32([1]), i32([2]), i32([10])
    n = np.array([10], np.int64)
    X, y = np.random.rand(20), np.random.rand(10)
    chunk = i32(np.arange(10))
    lvt, lnug = np.zeros(1), np.zeros(1)
    fixed = i32([0])
    P = ctypes.POINTER
    dp = lambda a: a.ctypes.data_as(P(ctypes.c_double))   # noqa: E731
    ip = lambda a: a.ctypes.data_as(P(ctypes.c_int32))    # noqa: E731
    rc = L.svh_batch_setup(gpu._h, 1, ip(d), ip(q), ip(nu), n.ctypes.data_as(P(ctypes.c_int64)),
                           dp(X), dp(y), None, None, None, ip(chunk), ip(nc), 5, dp(lvt),
                           ip(fixed), dp(lnug), 1e3)
    assert rc == 2
    assert b"Z is NULL" in L.svh_last_error(gpu._h)


def test_batch_rebuild_device_first_point(gpu):
    """Batched rebuilds with first = -1 (each dataset's first point found on
    the device) give the same structures as with the first points NumPy
    computes on the host (ordering.py:72): the objective of every dataset is
    bit-identical.  One dataset holds two rows at the same distance from the
    centroid (the lower index wins), one a duplicated row."""
    from pyscaledvecchia_amd import _lib
    from pyscaledvecchia_amd.gp import ScaledVecchiaGP
    from pyscaledvecchia_amd.likelihood import _chunk_ids_rows
    data = _datasets(11, nd=4)
    X0, y0 = data[0]
    c = X0.mean(0)
    X0[5] = c
    X0[len(X0) - 3] = c
    X1, y1 = data[1]
    X1[len(X1) // 2] = X1[3]
    models = [ScaledVecchiaGP(m_est=15, nu=2.5, trend="linear", n_est=2000) for _ in data]
    preps = [g._prepare(X, y) for g, (X, y) in zip(models, data)]
    nd = len(data)
    mmax = max(p["m"] for p in preps)
    chunks = [_chunk_ids_rows(len(p["ye"]), p["m"], p["d"] + 2) for p in preps]
    bh = _lib.BatchHandle(gpu)
    bh.setup([p["d"] for p in preps], [p["Ze"].shape[1] for p in preps], [2] * nd,
             [len(p["ye"]) for p in preps], None, None, None, None, None,
             np.concatenate([c for c, _ in chunks]), [nc for _, nc in chunks], mmax,
             [p["lvt"] for p in preps], [0] * nd, [0.0] * nd, models[0].lambda_max)
    bh.data(np.concatenate([p["Xe"].ravel() for p in preps]),
            np.concatenate([p["ye"] for p in preps]),
            np.concatenate([p["Ze"].ravel() for p in preps]))
    rng = np.random.default_rng(4)
    th = np.zeros((nd, 12))
    for i, p in enumerate(preps):
        th[i, :p["d"] + 2] = p["theta"] + 0.2 * rng.standard_normal(p["d"] + 2)
    lams = [np.exp(th[i, 1:1 + p["d"]]) for i, p in enumerate(preps)]
    which = list(range(nd))
    host = [O.maximin_first(p["Xe"] / la) for p, la in zip(preps, lams)]
    ms = [p["m"] for p in preps]
    bh.rebuild(which, lams, host, ms)
    r_host = bh.eval(th)
    bh.rebuild(which, lams, [-1] * nd, ms)
    r_dev = bh.eval(th)
    for a, b in zip(r_host, r_dev):
        np.testing.assert_array_equal(np.asarray(a), np.asarray(b))
