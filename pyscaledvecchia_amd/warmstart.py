"""Learned warm start for Fisher scoring.

No counterpart in the reference: its starting point (gp.py:169-177:
log var(y), log(0.2 sqrt d) for every range, log 0.1 for the nugget) stays
the default and the fallback.  Here a small network maps summary features of
a dataset to initial log-parameters (sigma^2, lambda_1..d, tau) in the
standardised space the fit works in; the fit then runs the same Fisher
scoring from that point.

Features (64, from a 2000-point subsample of the standardised data, its own
RNG, so the fit's RNG stream is untouched):
  * 60 = 12 dims x 5 lags: for the k-th nearest neighbour pairs
    (k = 1, 2, 4, 8, 16) the least-squares slopes s_l of the semivariogram
    0.5 (y_i - y_j)^2 ~ sum_l s_l (x_il - x_jl)^2, as log(max(s_l, 1e-8));
    dimensions beyond d are zero;
  * log of the mean semivariance at k = 1, excess kurtosis of y, nu / 2.5,
    d / 12.
Targets: the fitted log-parameters, padded to 14 = [sigma^2, lambda_1..12,
tau] with a mask of the d + 2 valid entries.  Features and targets are
normalised by their training means and standard deviations.  The starting
point takes sigma^2 from the network, log-ranges halfway between the
network's and the default start's, and the default nugget (see
unpack_theta).

Network: 64 -> 3 x (256, GELU) -> 14, bf16 weights with float32
accumulation; forward, backward and Adam are HIP kernels (svh_mlp_*).

Training data: datasets simulated from the prior of the batched benchmark
(X ~ U[0,1]^d stored as float32, nu in {0.5, 1.5, 2.5},
log10 lambda_l ~ U[-1, 1.5], tau ~ LogU[1e-8, 1e-2], sigma^2 = 1, d in
{2..10}: the batched fitter's kernels take d <= 10) by random Fourier
features on the GPU (svh_simulate_rff), fitted by the batched fitter.

    python -m pyscaledvecchia_amd.warmstart generate --datasets 20000 --out data.npz
    python -m pyscaledvecchia_amd.warmstart train --data data.npz --out net.npz
    python -m pyscaledvecchia_amd.warmstart evaluate --net net.npz --datasets 256 --n 50000
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np
from scipy.spatial import cKDTree

from . import _lib

N_IN, N_HIDDEN, N_LAYERS, N_OUT = 64, 256, 3, 14
D_MAX = 12
LAGS = (1, 2, 4, 8, 16)
N_SUB = 2000
DEFAULT_NET = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data",
                           "warmstart_net.npz")

__all__ = ["summary_features", "WarmStart", "simulate_prior", "build_training_set", "train"]


# ---------------------------------------------------------------- features
def summary_features(Xe, ye, nu, seed=0):
    """The 64 summary features of a standardised dataset (Xe in the unit box,
    ye centred and scaled as the fit standardises them)."""
    Xe = np.asarray(Xe, float)
    ye = np.asarray(ye, float)
    n, d = Xe.shape
    if d > D_MAX:
        raise ValueError(f"warm start supports d <= {D_MAX}")
    rng = np.random.default_rng(seed)
    idx = rng.choice(n, N_SUB, replace=False) if n > N_SUB else np.arange(n)
    X, y = Xe[idx], ye[idx]
    sd = y.std()
    y = (y - y.mean()) / (sd if sd > 0 else 1.0)
    k = min(max(LAGS), len(X) - 1)
    _, nb = cKDTree(X).query(X, k=k + 1)
    f = np.zeros(N_IN)
    g1 = 1.0
    for li, lag in enumerate(LAGS):
        j = nb[:, min(lag, k)]
        g = 0.5 * (y - y[j]) ** 2
        H = (X - X[j]) ** 2
        A = H.T @ H
        A += 1e-10 * (np.trace(A) / d + 1e-300) * np.eye(d)
        s = np.linalg.solve(A, H.T @ g)
        f[li * D_MAX: li * D_MAX + d] = np.log(np.maximum(s, 1e-8))
        if li == 0:
            g1 = max(float(g.mean()), 1e-12)
    kurt = float(np.mean(y ** 4) - 3.0)
    f[60:64] = [math.log(g1), kurt, nu / 2.5, d / D_MAX]
    return f.astype(np.float32)


def pack_targets(theta, d):
    """Fitted log-parameters -> (14,) target and mask."""
    t = np.zeros(N_OUT, np.float32)
    m = np.zeros(N_OUT, np.float32)
    t[0], m[0] = theta[0], 1
    t[1:1 + d], m[1:1 + d] = theta[1:1 + d], 1
    t[13], m[13] = theta[-1], 1
    return t, m


LOG_LAM_LO, LOG_LAM_HI = math.log(0.05), math.log(100.0)
LOG_NUG0 = math.log(0.1)


def unpack_theta(out, d):
    """Starting point from a network output: sigma^2 from the network, each
    log-range halfway between the network's and the default start's
    log(0.2 sqrt d) (then clipped to [0.05, 100]), the nugget from the default
    start (log 0.1).

    The nugget is barely predictable from the features (held-out correlation
    ~0.2) and small predicted nuggets sit where the likelihood is flat in tau,
    where Fisher scoring's stopping rule can end the search after one short
    step.  The ranges are damped because the objective has many stationary
    points (it changes with every reordering) and a range far beyond the unit
    box switches its input off, where the likelihood is flat in that range:
    with the network's ranges as they are, a sin-sum function of 8 inputs
    (n = 50,000, nugget ~3e-6) ends at half the default start's
    log-likelihood with one weak input switched off; halfway, it ends above
    it.  On 256 held-out prior draws (n = 50,000) the halfway start saves 1.25
    of 15.3 iterations (the undamped one 1.34), ends no worse than the
    default start for 64 % of them (61 %) and at worst 0.18 below in
    relative log-likelihood (0.35)."""
    lam = np.clip(0.5 * (out[1:1 + d] + math.log(0.2 * math.sqrt(d))), LOG_LAM_LO, LOG_LAM_HI)
    return np.concatenate([[out[0]], lam, [LOG_NUG0]]).astype(float)


# ---------------------------------------------------------------- network
class WarmStart:
    """Trained network plus normalisation; theta0() gives a starting point."""

    def __init__(self, params, fmean, fstd, tmean, tstd, device=0):
        self.params = np.asarray(params, np.float32)
        self.fmean, self.fstd = np.asarray(fmean, np.float32), np.asarray(fstd, np.float32)
        self.tmean, self.tstd = np.asarray(tmean, np.float32), np.asarray(tstd, np.float32)
        self.device = device
        self._h = None

    def _handle(self):
        if self._h is None:
            self._h = _lib.MlpHandle(_lib.context(self.device), N_IN, N_HIDDEN, N_LAYERS, N_OUT)
        self._h.set_params(self.params)
        return self._h

    def predict_raw(self, F):
        """Network output (unnormalised log-parameters, 14 wide) for features F."""
        Fn = (np.atleast_2d(F) - self.fmean) / self.fstd
        out = self._handle().forward(Fn)
        return out * self.tstd + self.tmean

    def theta0(self, Xe, ye, nu):
        return self.theta0_many([(Xe, ye, nu)])[0]

    def theta0_many(self, items):
        """Starting points for several standardised datasets [(Xe, ye, nu)]:
        features on host threads, one network forward for all."""
        from concurrent.futures import ThreadPoolExecutor
        if len(items) > 1:
            with ThreadPoolExecutor(max_workers=min(8, len(items))) as ex:
                F = list(ex.map(lambda t: summary_features(*t), items))
        else:
            F = [summary_features(*items[0])]
        out = self.predict_raw(np.array(F))
        return [unpack_theta(o, it[0].shape[1]) for o, it in zip(out, items)]

    def save(self, path):
        np.savez(path, params=self.params, fmean=self.fmean, fstd=self.fstd, tmean=self.tmean,
                 tstd=self.tstd, shape=np.array([N_IN, N_HIDDEN, N_LAYERS, N_OUT]))

    @classmethod
    def load(cls, path=None, device=0):
        z = np.load(path or DEFAULT_NET, allow_pickle=False)
        if tuple(z["shape"]) != (N_IN, N_HIDDEN, N_LAYERS, N_OUT):
            raise ValueError("warm-start network shape mismatch")
        return cls(z["params"], z["fmean"], z["fstd"], z["tmean"], z["tstd"], device=device)


# ---------------------------------------------------------------- data
def _spectral(rng, nu, d, F):
    """Frequencies of a unit-range Matern(nu) field: multivariate t with 2 nu
    degrees of freedom, omega = z sqrt(2 nu / u), u ~ chi2(2 nu)."""
    z = rng.standard_normal((F, d))
    u = rng.chisquare(2 * nu, (F, 1))
    return z * np.sqrt(2 * nu / u), rng.uniform(0, 2 * math.pi, F)


def simulate_prior(ctx, count, n, seed, d_range=(2, 10), F=1024, p_inert=0.0):
    """Datasets from the prior (X float32 storage, Matern nu draws by random
    Fourier features plus nugget noise).  With p_inert > 0 each input is, with
    that probability, inert (range 1e4: the draw is flat along it; at least
    one input stays active).  Returns [(X, y, nu, truth)]."""
    rng = np.random.default_rng(seed)
    specs, Xs, Ws, bs, amps = [], [], [], [], []
    for _ in range(count):
        d = int(rng.integers(d_range[0], d_range[1] + 1))
        nu = float(rng.choice([0.5, 1.5, 2.5]))
        lam = 10.0 ** rng.uniform(-1, 1.5, d)
        if p_inert > 0.0:
            inert = rng.random(d) < p_inert
            if inert.all():
                inert[int(rng.integers(d))] = False
            lam = np.where(inert, 1e4, lam)
        tau = math.exp(rng.uniform(math.log(1e-8), math.log(1e-2)))
        X = rng.random((n, d)).astype(np.float32).astype(np.float64)
        W, b = _spectral(rng, nu, d, F)
        specs.append((X, nu, lam, tau, rng.standard_normal(n)))
        Xs.append(X / lam)
        Ws.append(W)
        bs.append(b)
        amps.append(math.sqrt(2.0 / F))
    ys = []
    step = 512
    for a in range(0, count, step):
        ys += _lib.simulate_rff(ctx, Xs[a:a + step], Ws[a:a + step], bs[a:a + step],
                                amps[a:a + step])
    out = []
    for (X, nu, lam, tau, e), y in zip(specs, ys):
        out.append((X, y + math.sqrt(tau) * e, nu, dict(lam=lam, tau=tau)))
    return out


def build_training_set(count, n=5000, seed=1, batch=256, verbose=True, device=0,
                       checkpoint=None, p_inert=0.0):
    """Simulate `count` datasets, fit each with the batched fitter (default
    start, m_est = 30, constant trend), and return features, targets, masks.
    `checkpoint`: a path rewritten with the rows so far after each batch."""
    from .batch import BatchedFit
    from .gp import ScaledVecchiaGP
    ctx = _lib.context(device)
    F, T, M, info = [], [], [], []
    t0 = time.perf_counter()
    for a in range(0, count, batch):
        data = simulate_prior(ctx, min(batch, count - a), n, seed + a, p_inert=p_inert)
        models = [ScaledVecchiaGP(nu=nu, m_est=30, n_est=n, trend="constant",
                                  var_correction=False) for _, _, nu, _ in data]
        try:
            BatchedFit(models).fit([X for X, _, _, _ in data], [y for _, y, _, _ in data])
        except np.linalg.LinAlgError:
            # refit one by one and drop the failures
            keep = []
            for g, (X, y, nu, _) in zip(models, data):
                try:
                    BatchedFit([g]).fit([X], [y])
                    keep.append(True)
                except np.linalg.LinAlgError:
                    keep.append(False)
            models = [g for g, k in zip(models, keep) if k]
            data = [dd for dd, k in zip(data, keep) if k]
        for g, (X, y, nu, _) in zip(models, data):
            F.append(summary_features(g.X_, g.y_, nu))
            t, m = pack_targets(g.theta_, X.shape[1])
            T.append(t)
            M.append(m)
            info.append([X.shape[1], nu, g.loglik_, getattr(g, "n_iter_", -1), g.n_evals_])
        if verbose:
            print(f"  {a + batch}/{count} datasets fitted, {time.perf_counter() - t0:.1f} s",
                  flush=True)
        if checkpoint:
            np.savez(checkpoint, F=np.array(F), T=np.array(T), M=np.array(M),
                     info=np.array(info))
    return np.array(F), np.array(T), np.array(M), np.array(info)


def _init_params(seed):
    """He-uniform weights for the GELU layers, a smaller output layer, zero
    biases (float32, layout of svh_mlp_*)."""
    rng = np.random.default_rng(seed)
    dims = [N_IN] + [N_HIDDEN] * N_LAYERS + [N_OUT]
    parts = []
    for i in range(len(dims) - 1):
        a, b = dims[i], dims[i + 1]
        lim = math.sqrt(6.0 / a) if i < len(dims) - 2 else math.sqrt(1.0 / a)
        parts.append(rng.uniform(-lim, lim, a * b))
        parts.append(np.zeros(b))
    return np.concatenate(parts).astype(np.float32)


def train(F, T, M, epochs=60, batch=256, lr=1e-3, weight_decay=1e-3, seed=0, device=0,
          verbose=True, val_frac=0.1):
    """Train the network on the GPU (svh_mlp_* kernels).  Returns the
    WarmStart and a history of (epoch, train loss, validation loss)."""
    F = np.asarray(F, np.float32)
    T = np.asarray(T, np.float32)
    M = np.asarray(M, np.float32)
    rng = np.random.default_rng(seed)
    perm = rng.permutation(len(F))
    nval = int(round(val_frac * len(F)))
    va, tr = perm[:nval], perm[nval:]
    fmean, fstd = F[tr].mean(0), F[tr].std(0) + 1e-6
    msum = np.maximum(M[tr].sum(0), 1.0)
    tmean = (T[tr] * M[tr]).sum(0) / msum
    tstd = np.sqrt(((T[tr] - tmean) ** 2 * M[tr]).sum(0) / msum) + 1e-6
    Fn = (F - fmean) / fstd
    Tn = np.where(M > 0, (T - tmean) / tstd, 0.0)
    h = _lib.MlpHandle(_lib.context(device), N_IN, N_HIDDEN, N_LAYERS, N_OUT,
                       max_batch=max(batch, 1024))
    h.set_params(_init_params(seed))
    h.load_data(Fn[tr], Tn[tr], M[tr])
    hist = []
    best, best_p = float("inf"), None
    for ep in range(epochs):
        lr_ep = lr * 0.5 * (1 + math.cos(math.pi * ep / epochs))     # cosine schedule
        ltr = h.train_epoch(rng.permutation(len(tr)).astype(np.int32), batch, lr_ep,
                            weight_decay=weight_decay)
        lva = float("nan")
        if nval:
            out = h.forward(Fn[va])
            lva = float(((out - Tn[va]) ** 2 * M[va]).sum() / M[va].sum())
            if lva < best:                      # early stopping on the validation loss
                best, best_p = lva, h.get_params()
        hist.append((ep, ltr, lva))
        if verbose and (ep % 10 == 0 or ep == epochs - 1):
            print(f"  epoch {ep}: train {ltr:.4f}  val {lva:.4f}", flush=True)
    params = best_p if best_p is not None else h.get_params()
    ws = WarmStart(params, fmean, fstd, tmean, tstd, device=device)
    return ws, hist


# ---------------------------------------------------------------- evaluation
def evaluate(ws, count=256, n=50_000, seed=7, batch=64, device=0, verbose=True):
    """Held-out datasets fitted twice with the batched fitter: from the
    default start and from the learned one.  Returns per-dataset records."""
    from .batch import BatchedFit
    from .gp import ScaledVecchiaGP
    ctx = _lib.context(device)
    rec = []
    for a in range(0, count, batch):
        data = simulate_prior(ctx, min(batch, count - a), n, seed + 1_000_000 + a)
        runs = {}
        for mode in ("default", "learned"):
            models = [ScaledVecchiaGP(nu=nu, m_est=30, trend="constant", var_correction=False,
                                      warm_start=ws if mode == "learned" else None)
                      for _, _, nu, _ in data]
            t0 = time.perf_counter()
            BatchedFit(models).fit([X for X, _, _, _ in data], [y for _, y, _, _ in data])
            runs[mode] = (models, time.perf_counter() - t0)
        for i in range(len(data)):
            g0, g1 = runs["default"][0][i], runs["learned"][0][i]
            rec.append(dict(d=data[i][0].shape[1], nu=data[i][2],
                            iters_default=g0.n_iter_, iters_learned=g1.n_iter_,
                            evals_default=g0.n_evals_, evals_learned=g1.n_evals_,
                            loglik_default=g0.loglik_, loglik_learned=g1.loglik_,
                            learned_kept=bool(g1.warm_start_used_)))
        if verbose:
            print(f"  {a + len(data)}/{count}: batch time default "
                  f"{runs['default'][1]:.2f} s, learned {runs['learned'][1]:.2f} s", flush=True)
    return rec


def summarize(rec):
    it0 = np.array([r["iters_default"] for r in rec], float)
    it1 = np.array([r["iters_learned"] for r in rec], float)
    ev0 = np.array([r["evals_default"] for r in rec], float)
    ev1 = np.array([r["evals_learned"] for r in rec], float)
    l0 = np.array([r["loglik_default"] for r in rec])
    l1 = np.array([r["loglik_learned"] for r in rec])
    gap = (l1 - l0) / np.abs(l0)
    return dict(datasets=len(rec), iters_default_mean=float(it0.mean()),
                iters_learned_mean=float(it1.mean()),
                iters_saved_mean=float((it0 - it1).mean()),
                evals_default_mean=float(ev0.mean()), evals_learned_mean=float(ev1.mean()),
                loglik_rel_gap_median=float(np.median(gap)),
                loglik_rel_gap_min=float(gap.min()), loglik_rel_gap_max=float(gap.max()),
                frac_within_1e6=float(np.mean(np.abs(gap) <= 1e-6)),
                frac_learned_not_worse=float(np.mean(gap >= -1e-6)),
                frac_learned_kept=float(np.mean([r.get("learned_kept", True) for r in rec])))


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m pyscaledvecchia_amd.warmstart")
    sub = ap.add_subparsers(dest="cmd", required=True)
    g = sub.add_parser("generate")
    g.add_argument("--datasets", type=int, default=20000)
    g.add_argument("--n", type=int, default=5000)
    g.add_argument("--seed", type=int, default=1)
    g.add_argument("--batch", type=int, default=256)
    g.add_argument("--inert", type=float, default=0.0,
                   help="probability that an input of a simulated dataset is inert")
    g.add_argument("--out", required=True)
    t = sub.add_parser("train")
    t.add_argument("--data", required=True, nargs="+")
    t.add_argument("--out", required=True)
    t.add_argument("--epochs", type=int, default=60)
    t.add_argument("--batch", type=int, default=256)
    t.add_argument("--lr", type=float, default=1e-3)
    e = sub.add_parser("evaluate")
    e.add_argument("--net", default=None)
    e.add_argument("--datasets", type=int, default=256)
    e.add_argument("--n", type=int, default=50_000)
    e.add_argument("--batch", type=int, default=64)
    e.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    if a.cmd == "generate":
        t0 = time.perf_counter()
        F, T, M, info = build_training_set(a.datasets, a.n, a.seed, a.batch, checkpoint=a.out,
                                           p_inert=a.inert)
        np.savez(a.out, F=F, T=T, M=M, info=info)
        print(json.dumps({"datasets": len(F), "seconds": time.perf_counter() - t0}))
    elif a.cmd == "train":
        zs = [np.load(p, allow_pickle=False) for p in a.data]
        F = np.concatenate([z["F"] for z in zs])
        T = np.concatenate([z["T"] for z in zs])
        M = np.concatenate([z["M"] for z in zs])
        info = np.concatenate([z["info"] for z in zs])
        keep = info[:, 3] < 40          # fits that stopped before max_iter
        F, T, M = F[keep], T[keep], M[keep]
        t0 = time.perf_counter()
        ws, hist = train(F, T, M, epochs=a.epochs, batch=a.batch, lr=a.lr)
        ws.save(a.out)
        print(json.dumps({"rows": len(F), "seconds": time.perf_counter() - t0,
                          "final_train": hist[-1][1], "final_val": hist[-1][2],
                          "best_val": min(h[2] for h in hist)}))
    else:
        ws = WarmStart.load(a.net)
        rec = evaluate(ws, a.datasets, a.n, batch=a.batch)
        s = summarize(rec)
        if a.out:
            with open(a.out, "w") as f:
                json.dump(dict(summary=s, records=rec), f)
        print(json.dumps(s))
    return 0


if __name__ == "__main__":
    sys.exit(main())
