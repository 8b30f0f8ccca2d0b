"""Fit many independent datasets at once.

Each dataset is fitted exactly as `ScaledVecchiaGP.fit` fits it (reference
gp.py:138-239 with optimize.py:11-73): same standardisation, subsample,
starting values, reorder schedule and Fisher-scoring rules.  The difference is
where the work runs: every objective evaluation of every dataset that is
evaluating goes through one launch per kernel (uneven n, d, nu and trend are
mixed in one launch), the per-dataset Fisher-scoring state machine runs on the
GPU, and one step (evaluation + state update) is a captured hipGraph that the
host replays.  The host only steps in for the reorder hook (maximin order and
neighbours of the datasets whose ranges moved) and the final b-step.

Limits of the batched kernels: d <= 10, trend columns <= 11, m_est <= 31.
"""

from __future__ import annotations

import math
import time

import numpy as np

from . import _lib
from .gp import ScaledVecchiaGP
from .likelihood import _chunk_ids, _chunk_ids_rows
from .ordering import _nn_groups

__all__ = ["fit_batch", "BatchedFit"]

_NU_CODE = {0.5: 0, 1.5: 1, 2.5: 2}
_DONE, _FAILED, _REORDER = 3, 4, 2


class BatchedFit:
    """Batched fitting driver.

    models: ScaledVecchiaGP instances (one per dataset) holding the settings
    (m_est, n_est, nu, trend, nugget, var_penalty, lambda_max, max_iter, tol,
    random_state).  All models must share var_penalty, max_iter, tol and
    lambda_max (they parameterise the one device state machine).
    """

    def __init__(self, models, use_graph=True, device=0):
        self.models = list(models)
        self.use_graph = use_graph
        self.device = device
        m0 = self.models[0]
        for g in self.models:
            if g.nu not in _NU_CODE:
                raise ValueError("batched fits need nu in {0.5, 1.5, 2.5}")
            if (g.var_penalty, g.max_iter, g.tol, g.lambda_max) != \
                    (m0.var_penalty, m0.max_iter, m0.tol, m0.lambda_max):
                raise ValueError("batched models must share var_penalty, max_iter, tol, "
                                 "lambda_max")
        self.stats = {}

    def fit(self, Xs, ys):
        t_start = time.perf_counter()
        ctx = _lib.context(self.device)
        models = self.models
        nd = len(models)
        if len(Xs) != nd or len(ys) != nd:
            raise ValueError("one (X, y) per model")
        preps = []
        t0 = time.perf_counter()
        ws_idx = [i for i, g in enumerate(models) if g.theta0 is None and g.warm_start is not None]
        for i, (g, X, y) in enumerate(zip(models, Xs, ys)):
            if i in ws_idx:
                g._ws_theta0 = np.zeros(np.atleast_2d(X).shape[1] + 2)   # placeholder
            p = g._prepare(X, y)
            if p["d"] > 10 or g._design(np.zeros((1, p["d"]))).shape[1] > 11 or p["m"] > 31:
                raise ValueError("batched fits need d <= 10 and m_est <= 31")
            preps.append(p)
        if ws_idx:
            # learned starts of all warm-started datasets in one network call
            by_net = {}
            for i in ws_idx:
                by_net.setdefault(id(models[i]._warm_start_net()), []).append(i)
            for ids in by_net.values():
                net = models[ids[0]]._warm_start_net()
                ths = net.theta0_many([(preps[i]["Xe"], preps[i]["ye"], models[i].nu)
                                       for i in ids])
                for i, th in zip(ids, ths):
                    if models[i].nugget is not None:
                        th[-1] = math.log(models[i].nugget)
                    preps[i]["theta"] = models[i]._clip_theta(th, preps[i]["d"])
        # A learned start is judged by where its fit ends, not where it starts:
        # each warm-started dataset is fitted twice in the same batch, from
        # the learned start and from the reference start (gp.py:169-177), and
        # the fit with the higher final log-likelihood is kept (the reference
        # start on ties), so warm_start never returns a worse fit than the
        # reference start would.
        shadow = {i: nd + k for k, i in enumerate(ws_idx)}
        preps_all = preps + [dict(preps[i], theta=preps[i]["theta_default"].copy())
                             for i in ws_idx]
        models_all = models + [models[i] for i in ws_idx]
        na = len(preps_all)
        mmax = max(p["m"] for p in preps_all)
        # the batch buffers are allocated empty (no ordered data, no
        # neighbour lists); every structure is then built on the device
        d = np.array([p["d"] for p in preps_all], np.int32)
        q = np.array([p["Ze"].shape[1] for p in preps_all], np.int32)
        nu = np.array([_NU_CODE[g.nu] for g in models_all], np.int32)
        n = np.array([len(p["ye"]) for p in preps_all], np.int64)
        X = np.concatenate([np.ascontiguousarray(p["Xe"]).ravel() for p in preps_all])
        y = np.concatenate([p["ye"] for p in preps_all])
        Z = (np.concatenate([np.ascontiguousarray(p["Ze"]).ravel() for p in preps_all])
             if q.sum() else np.zeros(0))
        chunks = [_chunk_ids_rows(len(p["ye"]), p["m"], p["d"] + 2) for p in preps_all]
        chunk = np.concatenate([c for c, _ in chunks])
        nchunk = np.array([nc for _, nc in chunks], np.int32)
        lvt = np.array([p["lvt"] for p in preps_all])
        fixed = np.array([g.nugget is not None for g in models_all], np.int32)
        lnug = np.array([math.log(g.nugget) if g.nugget is not None else 0.0
                         for g in models_all])
        bh = _lib.BatchHandle(ctx)
        bh.setup(d, q, nu, n, None, None, None, None, None, chunk, nchunk, mmax, lvt, fixed,
                 lnug, models[0].lambda_max)
        bh.data(X, y, Z)
        self._rebuild(bh, preps_all, list(range(na)), [p["theta"] for p in preps_all])
        t_build = time.perf_counter() - t0
        theta0 = np.zeros((na, _lib.BP))
        for i, p in enumerate(preps_all):
            theta0[i, :p["d"] + 2] = p["theta"]
        last = [np.exp(p["theta"][1:1 + p["d"]]).copy() for p in preps_all]
        bh.fit_start(theta0, models[0].max_iter, models[0].tol, models[0].var_penalty,
                     self.use_graph)
        steps, rebuilds, t_reorder = 0, [1] * na, 0.0
        t0 = time.perf_counter()
        while True:
            st = bh.fit_step(1)
            steps += 1
            if np.all((st == _DONE) | (st == _FAILED)):
                break
            pend = np.flatnonzero(st == _REORDER)
            if pend.size:
                tr = time.perf_counter()
                th = bh.get()[0]
                redo = []
                for i in pend:
                    g, p = models_all[i], preps_all[i]
                    dd = p["d"]
                    thi = g._clip_theta(th[i, :dd + 2].copy(), dd)
                    r = np.exp(thi[1:1 + dd])
                    if np.max(np.abs(np.log(r / last[i]))) < 1e-3:
                        bh.set_reorder(i, False)
                        continue
                    redo.append(i)
                    last[i] = r.copy()
                if redo:
                    self._rebuild(bh, preps_all, redo,
                                  [th[i, :preps_all[i]["d"] + 2] for i in redo])
                    for i in redo:
                        bh.set_reorder(i, True)
                        rebuilds[i] += 1
                t_reorder += time.perf_counter() - tr
            if steps > 100000:
                raise RuntimeError("batched Fisher scoring did not terminate")
        t_loop = time.perf_counter() - t0
        th, ll, beta, st, iters, nev = bh.get()
        ws_won = 0
        for i, (g, p) in enumerate(zip(models, preps)):
            src = i
            if i in shadow:
                j = shadow[i]
                learned_ok = st[i] != _FAILED
                default_ok = st[j] != _FAILED
                if not (learned_ok and (not default_ok or ll[i] > ll[j])):
                    src = j
                g.warm_start_used_ = src == i
                ws_won += src == i
                g.n_evals_both_ = int(nev[i]) + int(nev[j])
            else:
                g.warm_start_used_ = False
            if st[src] == _FAILED:
                raise np.linalg.LinAlgError(f"dataset {i}: block Cholesky failed even with jitter")
            dd, qq = p["d"], int(q[i])
            g.n_evals_ = int(nev[src])
            g.n_rebuilds_ = rebuilds[src]
            g.n_iter_ = int(iters[src])
            g._finish(th[src, :dd + 2], float(ll[src]), beta[src, :qq].copy(), p["rng"], p["n"])
        self.stats = dict(steps=steps, wall_s=time.perf_counter() - t_start,
                          build_s=t_build, loop_s=t_loop, reorder_s=t_reorder,
                          evals=int(nev.sum()), rebuilds=int(sum(rebuilds)),
                          warm_started=len(ws_idx), learned_kept=int(ws_won))
        return models

    @staticmethod
    def _rebuild(bh, preps, which, thetas):
        """Device rebuild of datasets `which` at `thetas` (maximin order and
        ordered neighbours of the scaled inputs, gp.py:182-191); each first
        point (ordering.py:72) is found on the device (first = -1) with the
        NumPy arithmetic of the host computation."""
        lams = [np.exp(np.asarray(th, float)[1:1 + preps[i]["d"]])
                for i, th in zip(which, thetas)]
        ms = [preps[i]["m"] for i in which]
        bh.rebuild(which, lams, [-1] * len(which), ms)

    @staticmethod
    def _structure(g, p, theta, mmax):
        """Ordered inputs, per-row neighbour lists and jitter-chunk ids of one
        dataset at `theta` (gp.py:182-191 + likelihood.py:111-114)."""
        Xe, ye, Ze, m = p["Xe"], p["ye"], p["Ze"], p["m"]
        order, nn = g._order_structure(Xe, np.asarray(theta, float), m)
        groups = _nn_groups(nn, m)
        n = len(order)
        kcnt = np.minimum(np.arange(n), m).astype(np.int32)
        nbr = np.full((n, mmax), -1, np.int64)
        nbr[:, :m] = nn[:, 1:m + 1]
        chunk, nchunk = _chunk_ids(groups, p["d"] + 2)
        return dict(Xo=np.ascontiguousarray(Xe[order]), yo=np.ascontiguousarray(ye[order]),
                    Zo=np.ascontiguousarray(Ze[order]), nbr=nbr, kcnt=kcnt,
                    chunk=np.asarray(chunk, np.int32), nchunk=nchunk)


def fit_batch(Xs, ys, nus=None, use_graph=True, **kw):
    """Fit one ScaledVecchiaGP per (X, y) pair in one batched run.

    `nus` gives each dataset's smoothness (default: kw.get("nu", 2.5));
    other keyword arguments are ScaledVecchiaGP settings shared by all.
    Returns the fitted models (usable for predict / predict_joint).
    """
    nd = len(Xs)
    nus = [kw.pop("nu", 2.5)] * nd if nus is None else list(nus)
    models = [ScaledVecchiaGP(nu=nu, **kw) for nu in nus]
    bf = BatchedFit(models, use_graph=use_graph)
    bf.fit(Xs, ys)
    return models, bf.stats
