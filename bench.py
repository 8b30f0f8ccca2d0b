"""Benchmark: Vecchia loglik/grad/Fisher evaluations per second at n=50k, d=8, m=30.

Workload (one step = one objective evaluation: loglik + gradient (P=10) +
Fisher (10x10) + GLS beta (q=9)) on the n=50,000, d=8 synthetic test
function with a linear trend, nu=2.5, m_est=30, at the default starting
parameters of the fit and on the structure the fit builds first (maximin
order + ordered 30-NN in the scaled space).  Inputs are resident on the
device before the timed region.

Prints one JSON line.  Extra keys: the fit time of the same problem, the
rebuild (maximin + kNN) times, the dominant kernel's roofline, and a CPU
baseline (the C restatement of the oracle on the same structure).
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6          # MI355X FP64 vector / matrix
A_COEF = np.array([1.0, 0.8, 0.6, 0.4, 0.2, 0.05, 0.0, 0.0])
ELL = np.array([0.5, 0.7, 1.0, 1.5, 2.0, 3.0, 1.0, 1.0])


def c2_data(n=50_000, d=8, seed=1002):
    """X ~ U[0,1]^8; y = sum_l a_l sin(2 pi x_l / ell_l) + 0.1 x1 x2 + N(0, 1e-4)."""
    rng = np.random.default_rng(seed)
    X = rng.random((n, d))
    y = (A_COEF * np.sin(2 * np.pi * X / ELL)).sum(1) + 0.1 * X[:, 0] * X[:, 1]
    y = y + rng.normal(0.0, 0.01, n)
    return X, y


def c3_datasets(nd=64, seed=1003):
    """64 uneven datasets: n log-uniform in [1e3, 2e4], d uniform in {3..10},
    X ~ U[0,1]^d stored as float32, Matern nu in {0.5, 1.5, 2.5},
    log10 lambda_l ~ U[-1, 1.5], nugget tau ~ LogU[1e-8, 1e-2], sigma^2 = 1.
    y is an exact GP draw for n <= 3000, else an approximate draw by 2048
    random Fourier features of the same Matern spectral density."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(nd):
        n = int(round(math.exp(rng.uniform(math.log(1e3), math.log(2e4)))))
        d = int(rng.integers(3, 11))
        nu = float(rng.choice([0.5, 1.5, 2.5]))
        lam = 10.0 ** rng.uniform(-1, 1.5, d)
        tau = math.exp(rng.uniform(math.log(1e-8), math.log(1e-2)))
        X = rng.random((n, d)).astype(np.float32).astype(np.float64)
        Xs = X / lam
        if n <= 3000:
            from scipy.spatial.distance import cdist
            r = cdist(Xs, Xs)
            c = math.sqrt(2 * nu) * r
            M = {0.5: np.exp(-c), 1.5: (1 + c) * np.exp(-c),
                 2.5: (1 + c + c * c / 3) * np.exp(-c)}[nu]
            M[np.diag_indices(n)] += tau + 1e-10
            y = np.linalg.cholesky(M) @ rng.standard_normal(n)
        else:
            F = 2048
            # Matern spectral density: omega = z * sqrt(2 nu / chi2_{2nu}) (multivariate t)
            w = rng.standard_normal((F, d)) * np.sqrt(2 * nu / rng.chisquare(2 * nu, (F, 1)))
            b = rng.uniform(0, 2 * math.pi, F)
            y = math.sqrt(2.0 / F) * np.cos(Xs @ w.T + b).sum(1)
            y += math.sqrt(tau) * rng.standard_normal(n)
        out.append((X, y, nu))
    return out


def run_c3(args):
    from pyscaledvecchia_amd import ScaledVecchiaGP, _lib
    from pyscaledvecchia_amd.batch import BatchedFit
    data = c3_datasets(args.ndata)
    kw = dict(m_est=30, n_est=5000, trend="constant", var_correction=False)
    ctx = _lib.context()
    if not args.no_autotune:
        ctx.autotune(2)       # the batched pass 1's tile size, tuned by the warm-up batch
    # warm-up (kernels, allocations, the tile size of this batch shape)
    BatchedFit([ScaledVecchiaGP(nu=nu, **kw) for _, _, nu in data]).fit(
        [X for X, _, _ in data], [y for _, y, _ in data])
    models = [ScaledVecchiaGP(nu=nu, **kw) for _, _, nu in data]
    ctx.timing(True)
    ctx.timing_reset()
    bf = BatchedFit(models, use_graph=True)
    t0 = time.perf_counter()
    bf.fit([X for X, _, _ in data], [y for _, y, _ in data])
    t_batch = time.perf_counter() - t0
    ctx.timing(False)
    st_ms, st_n = ctx.timing_read("batch_step")
    out = {"metric": "batched fits/s, 64 uneven datasets (C3)", "value": len(data) / t_batch,
           "unit": "datasets/s", "n_gpus": 1, "higher_is_better": True, "data": "synthetic",
           "dtype": "f64", "config": {"workload": "C3 64 uneven datasets, m_est=30, n_est=5000",
                                      "datasets": len(data)},
           "batch_s": t_batch, "stats": bf.stats,
           "graph_step_ms_avg": st_ms / max(st_n, 1), "graph_steps": st_n,
           "autotune": {k: {"candidate": v[0], "ms": v[1]}
                        for k, v in ctx.autotune_choices().items()}}
    if not args.no_seq:
        t0 = time.perf_counter()
        seq = [ScaledVecchiaGP(nu=nu, **kw).fit(X, y) for X, y, nu in data]
        t_seq = time.perf_counter() - t0
        gaps = [abs(a.loglik_ - b.loglik_) / abs(b.loglik_) for a, b in zip(models, seq)]
        out["sequential_s"] = t_seq
        out["speedup_vs_sequential"] = t_seq / t_batch
        out["max_rel_loglik_gap_vs_sequential"] = max(gaps)
    print(json.dumps(out))


def run_c4(args):
    """Learned warm start on held-out datasets from the prior at n = 50,000:
    each fitted twice by the batched fitter, from the default start and from
    the network's start."""
    from pyscaledvecchia_amd import warmstart as W
    ws = W.WarmStart.load()
    t0 = time.perf_counter()
    rec = W.evaluate(ws, count=args.ndata, n=args.n_c4, batch=64, verbose=False)
    s = W.summarize(rec)
    out = {"metric": "warm start: Fisher-scoring iterations saved (mean), held-out datasets",
           "value": s["iters_saved_mean"], "unit": "iterations", "n_gpus": 1,
           "higher_is_better": True, "data": "synthetic (prior draws by random Fourier features)",
           "dtype": "f64 fit, bf16-weight network", "vs_baseline": None,
           "config": {"workload": f"C4 warm start, {args.ndata} held-out datasets, n={args.n_c4}",
                      "datasets": args.ndata, "n": args.n_c4},
           "summary": s, "seconds": time.perf_counter() - t0}
    print(json.dumps(out))


PEAK_FP64_MFMA_TFLOPS = 78.6


def run_mixed(args):
    """Mixed precision against FP64 per nugget decade: C2 inputs (n = 50,000,
    d = 8) with y = the sin-sum signal + N(0, tau var(signal)), tau = 1e-8 ..
    1e-2.  For each tau: one objective evaluation in both modes at the FP64
    estimate on one structure (loglik, gradient, Fisher, beta errors), whole
    fits in both modes (estimates compared raw and re-evaluated on one FP64
    structure), predictive mean and sd at 2000 test points, fit and pass-1
    times."""
    from pyscaledvecchia_amd import ScaledVecchiaGP, _lib
    ctx = _lib.context()
    n, d, m = args.n, 8, 30
    rng = np.random.default_rng(1002)
    X = rng.random((n, d))
    f = (A_COEF * np.sin(2 * np.pi * X / ELL)).sum(1) + 0.1 * X[:, 0] * X[:, 1]
    z = rng.standard_normal(n)
    Xt = np.random.default_rng(1012).random((2000, d))
    kw = dict(m_est=m, m_pred=140, nu=2.5, trend="linear", n_est=n)

    def rel(a, b):
        return float(np.max(np.abs(np.asarray(a) - b)) / np.max(np.abs(b)))
    ScaledVecchiaGP(**kw).fit(X, f + 0.01 * z)          # warm-up
    rows = []
    for tau in [1e-8, 1e-7, 1e-6, 1e-5, 1e-4, 1e-3, 1e-2]:
        y = f + math.sqrt(tau * f.var()) * z
        fits, times, p1 = {}, {}, {}
        for prec in ("fp64", "mixed", "mixed_f32cov"):
            ctx.timing(True)
            ctx.timing_reset()
            t0 = time.perf_counter()
            fits[prec] = ScaledVecchiaGP(precision=prec, **kw).fit(X, y)
            times[prec] = time.perf_counter() - t0
            ctx.timing(False)
            ms, cnt = ctx.timing_read("vecchia_pass1")
            p1[prec] = ms / max(cnt, 1)
        g6 = fits["fp64"]
        # one FP64 structure: the one resident after the last fit
        st = fits["mixed_f32cov"]._fit_state["struct"]
        lvt = math.log(max(g6.y_.var(), 1e-8))
        e64 = st(g6.theta_, 2.5, var_penalty=1.0, log_var_target=lvt)
        mu6, sd6 = g6.predict(Xt, return_std=True)
        row = {"tau": tau, "fitted_nugget": g6.nugget_, "fit_s": times, "pass1_ms": p1,
               "evals": {k: v.n_evals_ for k, v in fits.items()}}
        for prec in ("mixed", "mixed_f32cov"):
            gm = fits[prec]
            emx = st(g6.theta_, 2.5, var_penalty=1.0, log_var_target=lvt, precision=prec)
            ll_m_common = st(gm.theta_, 2.5, var_penalty=1.0, log_var_target=lvt)[0]
            mum, sdm = gm.predict(Xt, return_std=True)
            row[prec] = {
                "eval_rel_err": {"loglik": abs(emx[0] - e64[0]) / abs(e64[0]),
                                 "grad": rel(emx[1], e64[1]), "fisher": rel(emx[2], e64[2]),
                                 "beta": rel(emx[3], e64[3])},
                "fit_rel_err": {"loglik_raw": abs(gm.loglik_ - g6.loglik_) / abs(g6.loglik_),
                                "loglik_common_structure":
                                    abs(ll_m_common - e64[0]) / abs(e64[0]),
                                "theta": rel(gm.theta_, g6.theta_), "pred_mean": rel(mum, mu6),
                                "pred_sd": rel(sdm, sd6)},
                "refine_iters": gm.n_refine_}
        rows.append(row)
        print(json.dumps(rows[-1]), file=sys.stderr, flush=True)
    print(json.dumps({"metric": "mixed precision vs FP64 per nugget decade (C2 inputs)",
                      "unit": "relative error", "n_gpus": 1, "dtype": "f32+f64 vs f64",
                      "data": "synthetic (sin-sum signal, noise tau var)",
                      "config": {"workload": f"C2 inputs n={n} d={d} m_est={m} nu=2.5 linear"},
                      "rows": rows}))


def c5_inputs(n, ns, d=10, seed=1005):
    """Prediction workload after the large-prediction configuration: dims 1-3
    from a two-component Gaussian mixture (means 0.3 / 0.7, sd 0.1, clipped
    to [0, 1]), the rest uniform; y from the sin-sum test function extended
    to 10 inputs."""
    rng = np.random.default_rng(seed)

    def draw(k):
        X = rng.random((k, d))
        comp = rng.random(k) < 0.5
        mu = np.where(comp[:, None], 0.3, 0.7)
        X[:, :3] = np.clip(mu + 0.1 * rng.standard_normal((k, 3)), 0.0, 1.0)
        return X
    a = np.array([1.0, 0.8, 0.6, 0.4, 0.2, 0.05, 0.0, 0.0, 0.0, 0.0])[:d]
    ell = np.array([0.5, 0.7, 1.0, 1.5, 2.0, 3.0, 1.0, 1.0, 1.0, 1.0])[:d]
    X, Xs = draw(n), draw(ns)
    y = (a * np.sin(2 * np.pi * X / ell)).sum(1) + 0.1 * X[:, 0] * X[:, 1]
    return X, y + rng.normal(0.0, 0.01, n), Xs


def pred_block_flops(K, d):
    """Algorithmic flops of one prediction block (FMA = 2): pair distances and
    Matern, Cholesky K^3/3, the appended residual row, mean/variance."""
    T = K * (K + 1) / 2
    return T * (3 * d + 6) + K ** 3 / 3 + K ** 2 + 2 * K


def run_c5j(args):
    """Joint prediction and sampling on a path: sample_joint with n_sim = 100
    on n_test points after n_train observations (d = 10, m_pred = 140),
    predicted from a fixed fitted state; the solves run level by level on the
    device.  The CPU baseline times the reference's host solves
    (scipy spsolve_triangular) on the same factor."""
    from scipy import sparse
    from scipy.sparse.linalg import spsolve_triangular
    from pyscaledvecchia_amd import ScaledVecchiaGP, _lib
    ctx = _lib.context()
    d, m, nsim = 10, 140, 100
    X, y, Xs = c5_inputs(args.n_train, args.n_test, d)
    gp = ScaledVecchiaGP(m_est=30, m_pred=m, nu=2.5, trend="constant")
    # fitted state: standardised inputs and responses, fixed parameters
    gp._lo, gp._span = X.min(0), X.max(0) - X.min(0)
    gp._ymu, gp._ysd = y.mean(), y.std()
    gp.X_, gp.y_ = (X - gp._lo) / gp._span, (y - gp._ymu) / gp._ysd
    gp.d_ = d
    gp.theta_ = np.concatenate([[0.0], np.full(d, math.log(0.5)), [math.log(1e-4)]])
    gp.beta_ = np.array([0.0])
    gp.b_ = 1.0
    gp._estimate_nu = False
    gp._nu_fit = 2.5
    gp.predict_joint(Xs, n_sim=nsim, random_state=0)   # warm-up: device buffers at size
    ctx.timing(True)
    ctx.timing_reset()
    t0 = time.perf_counter()
    out = gp.predict_joint(Xs, n_sim=nsim, random_state=1)
    t_total = time.perf_counter() - t0
    ctx.timing(False)
    Xsn = (Xs - gp._lo) / gp._span
    t0 = time.perf_counter()
    block, av, perm, order_obs = gp._joint_factor(Xsn, m)
    t_factor = time.perf_counter() - t0
    ns, K, nobs = block.shape[0], block.shape[1], gp.X_.shape[0]
    resid = gp.y_ - gp.beta_[0]
    eps = np.random.default_rng(2).standard_normal((ns, nsim))
    t0 = time.perf_counter()
    mu, x, _ = ctx.joint_solve(nobs, block, av, resid=resid[order_obs], rhs=eps)
    t_solve = time.perf_counter() - t0
    # reference host solves on the same factor (CPU baseline, one core)
    U = sparse.coo_matrix((av.ravel(), (block.ravel(), np.repeat(np.arange(ns), K))),
                          shape=(nobs + ns, ns)).tocsc()
    Lt = U[nobs:, :].T.tocsr()
    t0 = time.perf_counter()
    ref_mu = spsolve_triangular(Lt, -(U[:nobs, :].T @ resid[order_obs]), lower=True)
    ref_x = spsolve_triangular(Lt, eps, lower=True)
    t_cpu = time.perf_counter() - t0
    rel = lambda a, b: float(np.max(np.abs(a - b)) / np.max(np.abs(b)))   # noqa: E731
    print(json.dumps({
        "metric": "sample_joint seconds, n_sim=100 on a path, m_pred=140, d=10",
        "value": t_total, "unit": "s", "n_gpus": 1, "steps": 1, "higher_is_better": False,
        "dtype": "f64", "vs_baseline": None, "data": "synthetic (GMM design, sin-sum function)",
        "config": {"workload": f"C5-shaped joint sampling n_train={args.n_train} "
                               f"path={args.n_test} n_sim={nsim} m_pred={m} d={d}"},
        "breakdown_s": {"factor (orderings, kNN, blocks)": t_factor,
                        "device solves (mean + 100 paths)": t_solve},
        "solve_vs_scipy_rel_err": {"mean": rel(mu, ref_mu), "paths": rel(x, ref_x)},
        "cpu_baseline": {"value": t_cpu, "unit": "s", "cores": 1, "kind": "reference",
                         "sample": "scipy spsolve_triangular (the reference's host solves, "
                                   "gp.py:380-401) on the same factor: mean + 100 paths"},
        "samples_shape": list(out["samples"].shape)}))


def run_c5p(args):
    """Marginal prediction at m_pred = 140 (K = 141 blocks): kNN query plus the
    matrix-core block kernel; inputs resident before the timed region."""
    from pyscaledvecchia_amd import _lib
    ctx = _lib.context()
    d, m = 10, 140
    X, y, Xs = c5_inputs(args.n_train, args.n_test, d)
    theta = np.concatenate([[0.0], np.full(d, math.log(0.5)), [math.log(1e-4)]])
    resid = (y - y.mean()) / y.std()
    ctx.predict(X, resid, Xs[:1000], m, theta, 2.5)           # warm-up
    res = {}
    for prec in ("fp64", "mixed"):
        ctx.timing(True)
        ctx.timing_reset()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            mu, va = ctx.predict(X, resid, Xs, m, theta, 2.5, precision=prec)
        t = (time.perf_counter() - t0) / args.steps
        ctx.timing(False)
        res[prec] = dict(wall_s=t, blocks_ms=ctx.timing_read("predict_blocks")[0] / args.steps,
                         knn_ms=ctx.timing_read("knn")[0] / args.steps, mu=mu, va=va)
    flops = args.n_test * pred_block_flops(m + 1, d)
    b64 = res["fp64"]["blocks_ms"] * 1e-3
    achieved = flops / b64 / 1e12
    rel = lambda a, b: float(np.max(np.abs(a - b)) / np.max(np.abs(b)))   # noqa: E731
    out = {"metric": "prediction blocks/s at m_pred=140 (K=141), d=10",
           "value": args.n_test / b64, "unit": "blocks/s", "n_gpus": 1, "steps": args.steps,
           "higher_is_better": True, "dtype": "f64", "vs_baseline": None,
           "data": "synthetic (GMM design, sin-sum function)",
           "config": {"workload": f"C5-shaped marginal prediction n_train={args.n_train} "
                                  f"n_test={args.n_test} d={d} m_pred={m}"},
           "roofline": {"bound": "mfma", "kernel": "pred_mfma_kernel<false>",
                        "achieved": achieved, "peak": PEAK_FP64_MFMA_TFLOPS, "unit": "TFLOP/s",
                        "frac": achieved / PEAK_FP64_MFMA_TFLOPS,
                        "flops_per_launch": flops, "kernel_avg_ms": 1e3 * b64},
           "knn_ms": res["fp64"]["knn_ms"], "predict_wall_s": res["fp64"]["wall_s"],
           "mixed": {"blocks_ms": res["mixed"]["blocks_ms"],
                     "mean_rel_err_vs_fp64": rel(res["mixed"]["mu"], res["fp64"]["mu"]),
                     "var_rel_err_vs_fp64": rel(res["mixed"]["va"], res["fp64"]["va"])}}
    print(json.dumps(out))


def block_flops(K, d, P, q):
    """Algorithmic FP64 flops of one block evaluation (FMA = 2): pairwise
    distance/Matern/g, dS.a for the d ranges, Cholesky, a and W solves, Fisher,
    GLS, z and gradient terms."""
    T = K * (K + 1) / 2
    return (T * (3 * d + 10 + 5 * d) + K ** 3 / 3 + (1 + P) * K ** 2 + P * (P + 1) * K
            + 2 * K * (1 + q) + q * (q + 1) + 2 * q + K ** 2 + 2 * K * q + 2 * P * K)


def traffic_file(kernel):
    return os.path.join(ROOT, "profiles", f"{kernel}_c2_traffic.json")


def calibrate_copy(nbytes=1 << 30):
    """A plain device-to-device copy of known size: the HBM counters
    (FETCH_SIZE, WRITE_SIZE) read for it give the scale between counter
    units and bytes."""
    import torch
    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes,
                      "check": float(b[-1].item())}))


def measured_traffic(kernel):
    """HBM bytes per launch of the named likelihood kernel from the committed
    counter runs (FETCH_SIZE + WRITE_SIZE, scaled by the copy calibration),
    or None when no such summary exists."""
    try:
        with open(traffic_file(kernel)) as f:
            return float(json.load(f)["hbm_bytes_per_launch"])
    except (OSError, KeyError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--n", type=int, default=50_000)
    ap.add_argument("--no-fit", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--workload", default="c2",
                    choices=["c2", "c3", "c4", "c5p", "c5j", "mixed"])
    ap.add_argument("--n-train", type=int, default=200_000)
    ap.add_argument("--n-test", type=int, default=100_000)
    ap.add_argument("--n-c4", type=int, default=50_000)
    ap.add_argument("--no-fast", action="store_true")
    ap.add_argument("--fit-only", action="store_true",
                    help="C2: one untimed and one timed fit, nothing else (for profilers)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="built-in launch schedules (default: tune during the warm-up)")
    ap.add_argument("--ndata", type=int, default=64)
    ap.add_argument("--no-seq", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 1 GiB device copy (HBM counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.workload == "c3":
        run_c3(args)
        return
    if args.workload == "mixed":
        run_mixed(args)
        return
    if args.workload == "c5j":
        if args.n_test == 100_000:
            args.n_test = 10_000     # the sampled path
        run_c5j(args)
        return
    if args.workload == "c5p":
        if args.steps == 50:
            args.steps = 5
        run_c5p(args)
        return
    if args.workload == "c4":
        if args.ndata == 64:
            args.ndata = 256
        run_c4(args)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)

    from pyscaledvecchia_amd import ScaledVecchiaGP, _lib
    from pyscaledvecchia_amd.likelihood import VecchiaStructure
    from pyscaledvecchia_amd.ordering import _nn_groups, find_ordered_nn, maximin_order

    n, d, m = args.n, 8, 30
    X, y = c2_data(n, d)
    if args.fit_only:
        kw = dict(m_est=m, m_pred=140, nu=2.5, trend="linear", n_est=n)
        ScaledVecchiaGP(**kw).fit(X, y)
        t0 = time.perf_counter()
        g = ScaledVecchiaGP(**kw).fit(X, y)
        print(json.dumps({"fit_s": time.perf_counter() - t0, "evals": g.n_evals_,
                          "rebuilds": g.n_rebuilds_, "loglik": g.loglik_}))
        return
    # the fit's own preprocessing (gp.py:149-177)
    lo, span = X.min(0), X.max(0) - X.min(0)
    Xe = (X - lo) / span
    ye = (y - y.mean()) / y.std()
    Ze = np.column_stack([np.ones(n), Xe])
    theta = np.concatenate([[math.log(ye.var())], np.full(d, math.log(0.2 * math.sqrt(d))),
                            [math.log(0.1)]])
    lvt = math.log(ye.var())
    ctx = _lib.context()
    if not args.no_autotune:
        ctx.autotune(2)       # schedules tuned by the first (warm-up) call of each shape

    # rebuild (timed separately, not part of the step)
    Xs = Xe / np.exp(theta[1:1 + d])
    maximin_order(Xs[:2000])
    ctx.timing(True)
    ctx.timing_reset()
    t0 = time.perf_counter()
    order = maximin_order(Xs)
    t_mm = time.perf_counter() - t0
    ctx.timing(False)
    mm_detail = {"rounds": ctx.timing_read("maximin_rounds")[1],
                 "apply_ms": ctx.timing_read("maximin_apply")[0],
                 "select_ms": ctx.timing_read("maximin_select")[0],
                 "gpu_span_ms": ctx.timing_read("maximin")[0]}
    t0 = time.perf_counter()
    nn = find_ordered_nn(Xs[order], m)
    t_knn = time.perf_counter() - t0
    groups = _nn_groups(nn, m)
    st = VecchiaStructure(Xe[order], ye[order], Ze[order], groups, ctx=ctx)
    P, q = d + 2, d + 1

    def timed_evals(precision):
        for _ in range(args.warmup):
            st(theta, 2.5, var_penalty=1.0, log_var_target=lvt, precision=precision)
        # the timed steps run without the library's event timers; the pass-1
        # kernel duration comes from a second run of the same steps with them
        t0 = time.perf_counter()
        for _ in range(args.steps):
            res = st(theta, 2.5, var_penalty=1.0, log_var_target=lvt, precision=precision)
        t = time.perf_counter() - t0     # every call synchronises on return
        ctx.timing(True)
        ctx.timing_reset()
        for _ in range(args.steps):
            st(theta, 2.5, var_penalty=1.0, log_var_target=lvt, precision=precision)
        ctx.timing(False)
        served = [k for k in ("ll_pass1_rows", "ll_pass1_split", "ll_pass1")
                  if ctx.timing_read("kernel:" + k)[1] > 0]
        return res, t, ctx.timing_read("vecchia_pass1"), served

    other = "mixed" if args.precision == "fp64" else "fp64"
    (ll, g, F, beta), t_steps, (k_ms, k_cnt), served = timed_evals(args.precision)
    (ll_o, g_o, F_o, _), t_o, (k_ms_o, k_cnt_o), served_o = timed_evals(other)
    kname = served[0] if len(served) == 1 else "+".join(served)
    ms_step = 1e3 * t_steps / args.steps
    g64, F64, ll64 = (g, F, ll) if args.precision == "fp64" else (g_o, F_o, ll_o)
    gmx, Fmx, llmx = (g_o, F_o, ll_o) if args.precision == "fp64" else (g, F, ll)

    def rel(a, b):
        return float(np.max(np.abs(np.asarray(a) - b)) / np.max(np.abs(b)))
    ks = np.minimum(np.arange(n), m) + 1
    flops = float(sum(block_flops(int(K), d, P, q) for K in ks[:m + 1]) +
                  (n - m - 1) * block_flops(m + 1, d, P, q))
    k_avg_s = 1e-3 * k_ms / max(k_cnt, 1)
    achieved = flops / k_avg_s / 1e12

    out = {
        "metric": "fit s + Vecchia loglik/grad/Fisher evals/s, n=50k d=8 m=30; FP32-mix vs FP64",
        "value": args.steps / t_steps,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64",
        "data": "synthetic (sin-sum test function, seeded)",
        "config": {"workload": f"C2 objective eval n={n} d={d} m_est={m} nu=2.5 trend=linear",
                   "global_batch": n, "seq_len": m + 1, "parallelism": "single-gpu",
                   "precision": args.precision},
        "roofline": {"bound": "mfma", "kernel": kname, "achieved": achieved,
                     "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS, "traffic": measured_traffic(kname),
                     "kernel_avg_ms": 1e3 * k_avg_s, "flops_per_launch": flops},
        "loglik": ll,
        other: {"evals_per_s": args.steps / t_o, "ms_per_step": 1e3 * t_o / args.steps,
                "kernel_avg_ms": k_ms_o / max(k_cnt_o, 1), "kernel": "+".join(served_o)},
        "mixed_vs_fp64_rel_err": {"loglik": abs(llmx - ll64) / abs(ll64),
                                  "grad": rel(gmx, g64), "fisher": rel(Fmx, F64)},
        "autotune": {k: {"candidate": v[0], "ms": v[1]} for k, v in ctx.autotune_choices().items()},
        "rebuild_s": {"maximin": t_mm, "ordered_knn": t_knn},
        "maximin_detail": mm_detail,
    }
    # full fit of the same problem (host control loop + GPU hot path)
    if not args.no_fit:
        Xt = np.random.default_rng(1012).random((2000, d))
        fits = {}
        # one untimed fit first: code objects loaded, device buffers at size
        ScaledVecchiaGP(m_est=m, m_pred=140, nu=2.5, trend="linear", n_est=n).fit(X, y)
        for prec in (args.precision, other):
            ctx.timing(True)
            ctx.timing_reset()
            t0 = time.perf_counter()
            gp = ScaledVecchiaGP(m_est=m, m_pred=140, nu=2.5, trend="linear", n_est=n,
                                 precision=prec).fit(X, y)
            t_fit = time.perf_counter() - t0
            ctx.timing(False)
            mu, sd = gp.predict(Xt, return_std=True)
            fits[prec] = (gp, mu, sd)
            if prec == "fp64":
                # evaluations per second at the fitted parameters, on the
                # fit's final structure (still resident)
                st_hat = gp._fit_state["struct"]
                kw_hat = dict(var_penalty=1.0, log_var_target=lvt)
                for _ in range(3):
                    st_hat(gp.theta_, 2.5, **kw_hat)
                t0 = time.perf_counter()
                for _ in range(args.steps):
                    st_hat(gp.theta_, 2.5, **kw_hat)
                t_hat = time.perf_counter() - t0
                ctx.timing(True)
                ctx.timing_reset()
                for _ in range(args.steps):
                    st_hat(gp.theta_, 2.5, **kw_hat)
                ctx.timing(False)
                ms_hat, cnt_hat = ctx.timing_read("vecchia_pass1")
                out["at_theta_hat"] = {"evals_per_s": args.steps / t_hat,
                                       "ms_per_step": 1e3 * t_hat / args.steps,
                                       "kernel_avg_ms": ms_hat / max(cnt_hat, 1),
                                       "theta_hat": gp.theta_.tolist()}
            out["fit_s" if prec == args.precision else f"fit_s_{prec}"] = t_fit
            out["fit" if prec == args.precision else f"fit_{prec}"] = {
                "evals": gp.n_evals_, "rebuilds": gp.n_rebuilds_,
                "refine_iters": getattr(gp, "n_refine_", 0),
                "loglik": gp.loglik_, "b": gp.b_,
                "maximin_s": ctx.timing_read("maximin")[0] / 1e3,
                "knn_s": ctx.timing_read("knn")[0] / 1e3,
                "pass1_s": ctx.timing_read("vecchia_pass1")[0] / 1e3,
                "predict_s": ctx.timing_read("predict_blocks")[0] / 1e3}
        (g6, mu6, sd6), (gm, mum, sdm) = fits["fp64"], fits["mixed"]
        out["mixed_fit_vs_fp64_rel_err"] = {
            "loglik": abs(gm.loglik_ - g6.loglik_) / abs(g6.loglik_),
            "theta": rel(gm.theta_, g6.theta_), "pred_mean": rel(mum, mu6),
            "pred_sd": rel(sdm, sd6), "test_points": len(Xt)}
        # fast paths: the batched fitter (device Fisher-scoring state machine
        # replayed as a hipGraph, device rebuilds) from the default start,
        # and the same with the learned warm start (start guard on)
        from pyscaledvecchia_amd.batch import BatchedFit
        from pyscaledvecchia_amd.warmstart import DEFAULT_NET
        legs = [("fast_fit", None)]
        if not args.no_fast and os.path.exists(DEFAULT_NET):
            legs.append(("fast_fit_learned", "learned"))
        for name, ws in legs:
            ctx.timing(True)
            ctx.timing_reset()
            t0 = time.perf_counter()
            gf = ScaledVecchiaGP(m_est=m, m_pred=140, nu=2.5, trend="linear", n_est=n,
                                 warm_start=ws)
            BatchedFit([gf]).fit([X], [y])
            t_fast = time.perf_counter() - t0
            ctx.timing(False)
            muf, sdf = gf.predict(Xt, return_std=True)
            out[f"{name}_s"] = t_fast
            out[name] = {
                "evals": gf.n_evals_, "rebuilds": gf.n_rebuilds_, "iters": gf.n_iter_,
                "iters_plain": g6.n_iter_, "warm_start_used": gf.warm_start_used_,
                "loglik": gf.loglik_, "speedup_vs_plain_fp64": out["fit_s"] / t_fast
                if args.precision == "fp64" else out["fit_s_fp64"] / t_fast,
                "maximin_s": ctx.timing_read("maximin")[0] / 1e3,
                "knn_s": ctx.timing_read("knn")[0] / 1e3,
                "vs_fp64_rel_err": {
                    "loglik": abs(gf.loglik_ - g6.loglik_) / abs(g6.loglik_),
                    "theta": rel(gf.theta_, g6.theta_), "pred_mean": rel(muf, mu6),
                    "pred_sd": rel(sdf, sd6)}}
    # CPU baseline: the C restatement of the oracle (OpenMP over blocks, as the
    # reference's prange), full C2 structure, repeated for ~cpu-seconds
    if not args.no_cpu:
        from oracle import c_oracle as C
        nthreads = max(1, min(16, len(os.sched_getaffinity(0))))
        Xo, yo, Zo = Xe[order], ye[order], Ze[order]
        reps, t0 = 0, time.perf_counter()
        while True:
            C.vecchia_profile_loglik(theta, Xo, yo, Zo, groups, 2.5, var_penalty=1.0,
                                     log_var_target=lvt, nthreads=nthreads)
            reps += 1
            if time.perf_counter() - t0 > args.cpu_seconds:
                break
        dt = (time.perf_counter() - t0) / reps
        out["cpu_baseline"] = {"value": 1.0 / dt, "unit": "evals/s", "cores": nthreads,
                               "kind": "port",
                               "sample": f"C oracle (liboracle_c.so, OpenMP over blocks), full "
                                         f"n={n} structure, {reps} evals"}
        # reference-faithful stage times: threads only over blocks (the
        # reference's prange sites) for the evaluation, serial maximin and
        # ordered kNN (ordering.py:27 is serial numba), each on a bounded
        # sample and scaled to n (maximin and brute-force kNN are O(n^2));
        # plus the NumPy oracle's evaluation on a slice, scaled by n
        from oracle import vecchia_np as ONP
        ns = min(n, 15_000)
        Xs_s = Xs[:ns]
        t0 = time.perf_counter()
        ord_s = C.maximin_order(Xs_s, int(np.argmin(((Xs_s - Xs_s.mean(0)) ** 2).sum(1))))
        t_mm_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        C.find_ordered_nn(Xs_s[ord_s], m)
        t_knn_s = time.perf_counter() - t0
        scale2 = (n / ns) ** 2
        n_np = 5000
        g_np = _nn_groups(nn[:n_np], m)
        t0 = time.perf_counter()
        ONP.vecchia_profile_loglik(theta, Xo[:n_np], yo[:n_np], Zo[:n_np], g_np, 2.5,
                                   var_penalty=1.0, log_var_target=lvt)
        t_np = (time.perf_counter() - t0) * n / n_np
        evals_fit = out.get("fit", {}).get("evals", 48)
        rebuilds_fit = out.get("fit", {}).get("rebuilds", 5)
        out["cpu_baseline"]["reference_faithful"] = {
            "eval_s": dt, "eval_threads": nthreads,
            "maximin_s_serial": t_mm_s * scale2, "ordered_knn_s_serial": t_knn_s * scale2,
            "sample": f"maximin / ordered kNN timed at n={ns} (serial C), scaled by (n/{ns})^2",
            "fit_s_estimate": rebuilds_fit * (t_mm_s + t_knn_s) * scale2 + evals_fit * dt,
            "fit_estimate_rule": f"{rebuilds_fit} rebuilds x (maximin + kNN) + {evals_fit} evals "
                                 "(the GPU fit's counts)"}
        out["cpu_baseline"]["numpy_oracle_eval_s"] = {
            "value": t_np, "sample": f"oracle/vecchia_np.py on the first {n_np} rows, "
                                     f"scaled by n/{n_np}"}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
