"""NumPy FP64 restatement of the reference scaled-Vecchia path (oracle).

TEST INFRASTRUCTURE ONLY -- imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg.  Never imported by the product package.

Reference: jdtuck/pyscaledvecchia, src/scaled_vecchia/{ordering,_covariance,
likelihood,optimize,gp}.py.  Every function below cites the file:line range
whose behaviour it restates.

Canonical arithmetic (the reference runs these loops under numba fastmath,
whose summation order is host dependent):
  * a squared distance is the sequential sum over l of diff_l*diff_l, each
    product rounded before the add (no FMA), l = 0..d-1;
  * neighbour sets and within-row order are by (squared distance, index);
  * the maximin arg-max takes the lowest index among equal distances.
The HIP kernels follow the same definitions, so ordering and neighbour sets
compare bit-exactly.
"""

from __future__ import annotations

import math

import numpy as np
import scipy.linalg
import scipy.special
from scipy import sparse
from scipy.sparse.linalg import spsolve_triangular

LOG2PI = math.log(2.0 * math.pi)          # _covariance.py:39
SQRT3 = math.sqrt(3.0)
SQRT5 = math.sqrt(5.0)
CHUNK_BUDGET = 4_000_000                  # _covariance.py:255
JITTER_POWERS = tuple(range(-12, -3))     # _covariance.py:247 (p = -12 .. -4)


# --------------------------------------------------------------------------
# distances (canonical rounding)
# --------------------------------------------------------------------------

def sqdist_to(X, x):
    """Squared distances from every row of X to x: sequential sum over l."""
    X = np.asarray(X, dtype=np.float64)
    s = np.zeros(X.shape[0])
    for l in range(X.shape[1]):
        t = X[:, l] - x[l]
        s = s + t * t
    return s


# --------------------------------------------------------------------------
# Matern closed forms (_covariance.py:49-72)
# --------------------------------------------------------------------------

def matern_corr(r, nu):
    """M_nu(r) for nu in {0.5, 1.5, 2.5} (_covariance.py:49-59)."""
    r = np.asarray(r, dtype=np.float64)
    if nu == 0.5:
        return np.exp(-r)
    if nu == 1.5:
        c = SQRT3 * r
        return (1.0 + c) * np.exp(-c)
    if nu == 2.5:
        c = SQRT5 * r
        return (1.0 + c + c * c / 3.0) * np.exp(-c)
    return matern_corr_general(r, nu)     # _covariance.py:178-183 falls back


def matern_g(r, nu):
    """g(r) = M'(r)/r; 0 at r=0 for nu=0.5 (_covariance.py:62-72, 125-130)."""
    r = np.asarray(r, dtype=np.float64)
    if nu == 0.5:
        rs = np.where(r > 0.0, r, 1.0)
        return np.where(r > 0.0, -np.exp(-rs) / rs, 0.0)
    if nu == 1.5:
        return -3.0 * np.exp(-SQRT3 * r)
    if nu == 2.5:
        c = SQRT5 * r
        return -(5.0 / 3.0) * (1.0 + c) * np.exp(-c)
    return matern_g_general(r, nu)


R_FLOOR = 1e-7     # _matern_general.py:88


def matern_corr_general(r, nu):
    """M_nu(r) = 2^(1-nu)/Gamma(nu) c^nu K_nu(c), c = sqrt(2 nu) r; 1 below
    the distance floor (_matern_general.py:92-99)."""
    r = np.asarray(r, dtype=np.float64)
    rr = np.where(r < R_FLOOR, R_FLOOR, r)
    c = math.sqrt(2.0 * nu) * rr
    lc = (1.0 - nu) * math.log(2.0) - scipy.special.gammaln(nu)
    val = np.exp(lc + nu * np.log(c)) * scipy.special.kv(nu, c)
    return np.where(r < R_FLOOR, 1.0, val)


def matern_g_general(r, nu):
    """g(r) = M_nu'(r)/r via d/dz[z^nu K_nu(z)] = -z^nu K_{nu-1}(z); 0 below
    the floor (_matern_general.py:102-116)."""
    r = np.asarray(r, dtype=np.float64)
    rr = np.where(r < R_FLOOR, R_FLOOR, r)
    c = math.sqrt(2.0 * nu) * rr
    lc = (1.0 - nu) * math.log(2.0) - scipy.special.gammaln(nu)
    val = -(math.sqrt(2.0 * nu) / rr) * np.exp(lc + nu * np.log(c)) * scipy.special.kv(nu - 1.0, c)
    return np.where(r < R_FLOOR, 0.0, val)


# --------------------------------------------------------------------------
# ordering (ordering.py:27-127)
# --------------------------------------------------------------------------

def maximin_first(X):
    """First point: closest to the centroid (ordering.py:71-72)."""
    X = np.ascontiguousarray(X, dtype=np.float64)
    return int(np.argmin(((X - X.mean(0)) ** 2).sum(1)))


def maximin_order(X, first=None, steps=None):
    """Exact maximin order, serial O(n^2 d) (ordering.py:27-73).

    Step t picks the unselected point with the largest distance to the
    selected set (lowest index among ties, ordering.py:43-47) and lowers every
    unselected distance that the new point beats strictly (ordering.py:48-57).
    `steps` (test aid) stops after that many points and returns the prefix.
    """
    X = np.ascontiguousarray(X, dtype=np.float64)
    n = X.shape[0]
    if first is None:
        first = maximin_first(X)
    order = np.empty(n, dtype=np.int64)
    order[0] = first
    if n == 0:
        return order[:0]
    dist = np.sqrt(sqdist_to(X, X[first]))
    dist[first] = -1.0
    n_out = n if steps is None else min(n, steps)
    for t in range(1, n_out):
        j = int(np.argmax(dist))             # first maximum = lowest index
        order[t] = j
        dj = np.sqrt(sqdist_to(X, X[j]))
        upd = (dist >= 0.0) & (dj < dist)
        dist[upd] = dj[upd]
        dist[j] = -1.0
    return order[:n_out]


def find_ordered_nn(X, m, start=0):
    """Ordered nearest neighbours by brute force (ordering.py:76-116).

    Row i: column 0 is i, columns 1..min(i,m) the min(i,m) nearest j < i by
    (squared distance, index), nearest first; -1 elsewhere.  Rows below
    `start` (other than column 0) stay -1, as in the reference.
    """
    X = np.ascontiguousarray(X, dtype=np.float64)
    n = X.shape[0]
    nn = np.full((n, m + 1), -1, dtype=np.int64)
    nn[:, 0] = np.arange(n)
    for i in range(max(start, 1), n):
        k = min(i, m)
        s = sqdist_to(X[:i], X[i])
        o = np.lexsort((np.arange(i), s))[:k]
        nn[i, 1:k + 1] = o
    return nn


def nn_groups(nn, m):
    """Rows grouped by conditioning-set size (ordering.py:119-127)."""
    n = nn.shape[0]
    ks = np.minimum(np.arange(n), m)
    groups = []
    for k in np.unique(ks):
        rows = np.where(ks == k)[0]
        groups.append((rows, nn[rows, 1:k + 1]))
    return groups


def knn_query(Xref, Xq, k):
    """k nearest rows of Xref for every row of Xq, sorted by (sq. distance,
    index).  Restates the cKDTree query of gp.py:253-254 (sorted output)."""
    Xref = np.ascontiguousarray(Xref, dtype=np.float64)
    Xq = np.ascontiguousarray(Xq, dtype=np.float64)
    out = np.empty((Xq.shape[0], k), dtype=np.int64)
    idx = np.arange(Xref.shape[0])
    for i in range(Xq.shape[0]):
        s = sqdist_to(Xref, Xq[i])
        out[i] = np.lexsort((idx, s))[:k]
    return out


# --------------------------------------------------------------------------
# block covariance, Cholesky, triangular solves (_covariance.py:86-259)
# --------------------------------------------------------------------------

def block_cov(Xb, ranges, variance, nugget, nu, derivs=True, nug_mask=None,
              estimate_nu=False, nu_fd_h=1e-3):
    """Sigma (B,K,K) and dSigma/dlog-theta (B,P,K,K), P = d+2 (d+3 with
    estimate_nu: the log nu column, nu * central difference of M in nu with
    step 1e-3, sits before the nugget's, _matern_general.py:134-179).

    _covariance.py:86-147 (kernels) and 150-198 (dispatch).  Sigma_ij =
    variance*(M(r_ij) + nugget*mask_i*[i==j]); dS_0 = Sigma; dS_{1+l} =
    -variance*g(r)*((x_il-x_jl)/lambda_l)^2; dS_{1+d} = variance*nugget*mask*I.
    """
    Xb = np.asarray(Xb, dtype=np.float64)
    ranges = np.asarray(ranges, dtype=np.float64)
    B, K, d = Xb.shape
    diff = (Xb[:, :, None, :] - Xb[:, None, :, :]) / ranges
    s = np.zeros((B, K, K))
    for l in range(d):
        s = s + diff[..., l] * diff[..., l]
    r = np.sqrt(s)
    M = matern_corr(r, nu)
    mask = np.ones((B, K)) if nug_mask is None else np.asarray(nug_mask, dtype=np.float64)
    nug = nugget * mask[:, :, None] * np.eye(K)[None]
    Sigma = variance * (M + nug)
    if not derivs:
        return Sigma, None
    g = matern_g(r, nu)
    P = d + 2 + (1 if estimate_nu else 0)
    dS = np.empty((B, P, K, K))
    dS[:, 0] = Sigma
    for l in range(d):
        dS[:, 1 + l] = -variance * g * diff[..., l] * diff[..., l]
    if estimate_nu:
        dMnu = (matern_corr_general(r, nu + nu_fd_h)
                - matern_corr_general(r, nu - nu_fd_h)) / (2.0 * nu_fd_h)
        dS[:, 1 + d] = variance * nu * dMnu
    dS[:, P - 1] = variance * nug
    return Sigma, dS


class CholeskyFailed(np.linalg.LinAlgError):
    pass


def batch_chol(Sigma):
    """Chunk-level jitter ladder (_covariance.py:235-252): if any block of the
    chunk fails, the whole chunk is retried with +10^p*tr(S_b)/K*I,
    p = -12..-4; after that LinAlgError."""
    try:
        return np.linalg.cholesky(Sigma), None
    except np.linalg.LinAlgError:
        pass
    K = Sigma.shape[-1]
    scale = np.einsum("bii->b", Sigma) / K
    eye = np.eye(K)
    for p in JITTER_POWERS:
        try:
            return np.linalg.cholesky(Sigma + (10.0 ** p) * scale[:, None, None] * eye), p
        except np.linalg.LinAlgError:
            continue
    raise CholeskyFailed("block Cholesky failed even with jitter")


def forward_solve(L, Bm):
    """L X = B for lower-triangular L, batched (_covariance.py:205-217)."""
    b, K, nrhs = Bm.shape
    X = np.empty_like(Bm, dtype=np.float64)
    for i in range(K):
        acc = Bm[:, i, :].astype(np.float64).copy()
        if i:
            acc -= np.einsum("bj,bjr->br", L[:, i, :i], X[:, :i, :])
        X[:, i, :] = acc / L[:, i, i][:, None]
    return X


def backsolve_LT(L, Bm):
    """L^T X = B for lower-triangular L, batched (_covariance.py:220-232)."""
    b, K, nrhs = Bm.shape
    X = np.empty_like(Bm, dtype=np.float64)
    for i in range(K - 1, -1, -1):
        acc = Bm[:, i, :].astype(np.float64).copy()
        if i < K - 1:
            acc -= np.einsum("bj,bjr->br", L[:, i + 1:, i], X[:, i + 1:, :])
        X[:, i, :] = acc / L[:, i, i][:, None]
    return X


def chunk_step(K, P, budget=CHUNK_BUDGET):
    """Rows per chunk (_covariance.py:255-259)."""
    return max(1, budget // max(1, K * K * max(P, 1)))


def chunks(n_rows, K, P, budget=CHUNK_BUDGET):
    step = chunk_step(K, P, budget)
    for a in range(0, n_rows, step):
        yield a, min(n_rows, a + step)


# --------------------------------------------------------------------------
# profile Vecchia loglik / gradient / Fisher (likelihood.py:66-175)
# --------------------------------------------------------------------------

def vecchia_profile_loglik(theta, Xo, yo, Zo, groups, nu, need_grad=True,
                           var_penalty=0.0, log_var_target=0.0, report=None,
                           estimate_nu=False):
    """Two-pass evaluation of likelihood.py:66-175 (fixed nu only).

    Pass 1 per chunk: Sigma/dS, Cholesky, a = L^-T e_K, GLS normal equations,
    W = L^-1 (dS a), Fisher.  beta = solve(XtX + 1e-12 I, Zty).  Pass 2:
    z = L^-1 (y - Z beta), loglik and gradient.  Optional sigma^2 penalty.
    `report`, if a dict, receives the jitter power used per chunk.
    """
    theta = np.asarray(theta, dtype=np.float64)
    d = Xo.shape[1]
    variance = math.exp(theta[0])
    ranges = np.exp(theta[1:1 + d])
    if estimate_nu:                           # likelihood.py:90-93
        nu = math.exp(theta[1 + d])
        nugget = math.exp(theta[2 + d])
        P = d + 3
    else:
        nugget = math.exp(theta[1 + d])
        P = d + 2
    n = Xo.shape[0]
    has_Z = Zo is not None and Zo.shape[1] > 0
    q = Zo.shape[1] if has_Z else 0
    XtX = np.zeros((q, q))
    Zty = np.zeros(q)
    fish = np.zeros((P, P))
    cache = []
    jit = []
    for rows, nb in groups:
        K = nb.shape[1] + 1
        idx = np.concatenate([nb, rows[:, None]], axis=1)
        for a0, a1 in chunks(len(rows), K, P if need_grad else 1):
            block = idx[a0:a1]
            Sig, dS = block_cov(Xo[block], ranges, variance, nugget, nu, derivs=need_grad,
                                estimate_nu=estimate_nu)
            L, p = batch_chol(Sig)
            jit.append(p)
            E = np.zeros((a1 - a0, K, 1))
            E[:, K - 1, 0] = 1.0
            av = backsolve_LT(L, E)[:, :, 0]
            if has_Z:
                V = np.concatenate([yo[block][:, :, None], Zo[block]], axis=2)
                Vt = np.einsum("bk,bkq->bq", av, V)
                yt, Zt = Vt[:, 0], Vt[:, 1:]
                XtX += Zt.T @ Zt
                Zty += Zt.T @ yt
            W = None
            if need_grad:
                rhs = np.einsum("bpkl,bl->bkp", dS, av)
                W = forward_solve(L, rhs).transpose(0, 2, 1)
                wK = W[:, :, K - 1]
                fish += 0.5 * (2.0 * np.einsum("bpk,bqk->pq", W, W) - wK.T @ wK)
            cache.append((block, K, L, W))
    beta = np.linalg.solve(XtX + 1e-12 * np.eye(q), Zty) if has_Z else np.zeros(0)
    ll = -0.5 * n * LOG2PI
    grad = np.zeros(P)
    for block, K, L, W in cache:
        resid = (yo[block] - Zo[block] @ beta) if has_Z else yo[block]
        z = forward_solve(L, resid[:, :, None])[:, :, 0]
        zK = z[:, K - 1]
        ll -= 0.5 * (2.0 * np.log(L[:, K - 1, K - 1]) + zK ** 2).sum()
        if not need_grad:
            continue
        wK = W[:, :, K - 1]
        wz = np.einsum("bpk,bk->bp", W, z)
        grad -= 0.5 * (wK * (1.0 + zK[:, None] ** 2) - 2.0 * zK[:, None] * wz).sum(0)
    if var_penalty > 0.0:
        excess = max(0.0, theta[0] - log_var_target)
        ll -= 0.5 * var_penalty * excess ** 2
        if need_grad:
            grad[0] -= var_penalty * excess
            if excess > 0.0:
                fish[0, 0] += var_penalty
    if isinstance(report, dict):
        report["jitter"] = jit
    if not need_grad:
        return ll, None, None, beta
    return ll, grad, fish, beta


# --------------------------------------------------------------------------
# Fisher scoring (optimize.py:11-73)
# --------------------------------------------------------------------------

def fisher_scoring(theta0, objective, max_iter=40, tol=1e-4, reorder_hook=None,
                   max_step=3.0, trace=None):
    """Ridge-until-ascent Fisher step, step cap, strict-increase backtracking,
    reorder hook at k = 2, 4, 8, ... (optimize.py:11-73)."""
    theta = np.array(theta0, dtype=np.float64)
    ll, g, M, beta = objective(theta)
    next_reorder = 2
    for k in range(1, max_iter + 1):
        if reorder_hook is not None and k == next_reorder:
            next_reorder *= 2
            if reorder_hook(theta, k):
                ll, g, M, beta = objective(theta)
        step = None
        ridge = 0.0
        diag = np.maximum(np.diag(M), 1e-8)
        for _ in range(12):
            try:
                cand = np.linalg.solve(M + ridge * np.diag(diag), g)
            except np.linalg.LinAlgError:
                ridge = max(1e-6, ridge * 10.0)
                continue
            if np.dot(cand, g) > 0 and np.all(np.isfinite(cand)):
                step = cand
                break
            ridge = max(1e-6, ridge * 10.0)
        if step is None:
            step = g / max(np.linalg.norm(g), 1e-12)
        nrm = np.linalg.norm(step)
        if nrm > max_step:
            step *= max_step / nrm
        alpha, ok = 1.0, False
        for _ in range(20):
            try:
                c_ll, c_g, c_M, c_beta = objective(theta + alpha * step)
            except np.linalg.LinAlgError:
                alpha *= 0.5
                continue
            if np.isfinite(c_ll) and c_ll > ll:
                theta = theta + alpha * step
                ll, g, M, beta = c_ll, c_g, c_M, c_beta
                ok = True
                break
            alpha *= 0.5
        if trace is not None:
            trace.append((k, float(ll)))
        if not ok:
            break
        if abs(np.dot(alpha * step, g)) < tol:
            break
    return theta, ll, beta


# --------------------------------------------------------------------------
# estimator pipeline (gp.py:94-425)
# --------------------------------------------------------------------------

def design(X, trend):
    """gp.py:94-101."""
    if trend == "zero":
        return np.zeros((X.shape[0], 0))
    if trend == "constant":
        return np.ones((X.shape[0], 1))
    if trend == "linear":
        return np.column_stack([np.ones(X.shape[0]), X])
    raise ValueError("trend must be 'zero', 'constant' or 'linear'")


def clip_theta(theta, d, lambda_max=1e3, estimate_nu=False):
    """gp.py:126-135."""
    hi = math.log(lambda_max) if np.isfinite(lambda_max) else np.inf
    theta[1:1 + d] = np.clip(theta[1:1 + d], math.log(1e-8), hi)
    if estimate_nu:
        theta[1 + d] = np.clip(theta[1 + d], math.log(0.3), math.log(10.0))
    theta[-1] = np.clip(theta[-1], math.log(1e-12), math.log(1e3))
    return theta


class OracleGP:
    """Restatement of ScaledVecchiaGP.fit/predict/predict_joint (gp.py:40-425)
    for fixed nu in {0.5, 1.5, 2.5}."""

    def __init__(self, m_est=30, m_pred=140, n_est=5000, nu=2.5, trend="constant",
                 nugget=None, var_penalty=1.0, var_correction=True,
                 lambda_max=1e3, max_iter=40, tol=1e-4, random_state=0):
        self.m_est, self.m_pred, self.n_est, self.nu = m_est, m_pred, n_est, nu
        self._nu_fit = nu
        self.trend, self.nugget, self.var_penalty = trend, nugget, var_penalty
        self.var_correction, self.lambda_max = var_correction, lambda_max
        self.max_iter, self.tol, self.random_state = max_iter, tol, random_state
        self.trace = []
        self.n_evals = 0
        self.n_rebuilds = 0

    @property
    def ranges_(self):
        return np.exp(self.theta_[1:1 + self.d_])

    @property
    def nu_(self):
        return self._nu_fit

    def fit(self, X, y):
        """gp.py:138-239."""
        X = np.ascontiguousarray(np.atleast_2d(X), dtype=float)
        y = np.asarray(y, dtype=float).ravel()
        n, d = X.shape
        if y.shape[0] != n:
            raise ValueError("X and y have incompatible shapes")
        self.d_ = d
        self._lo = X.min(0)
        self._span = np.where(X.max(0) - X.min(0) > 0, X.max(0) - X.min(0), 1.0)
        self._ymu = y.mean()
        self._ysd = y.std() if y.std() > 0 else 1.0
        self.X_ = (X - self._lo) / self._span
        self.y_ = (y - self._ymu) / self._ysd
        rng = np.random.default_rng(self.random_state)
        if self.n_est is not None and n > self.n_est:
            sub = rng.choice(n, self.n_est, replace=False)
        else:
            sub = np.arange(n)
        Xe, ye = self.X_[sub], self.y_[sub]
        Ze = design(Xe, self.trend)
        m = min(self.m_est, len(sub) - 1)
        est = self.nu is None                     # gp.py:145, 169-177
        self._estimate_nu = est
        theta = np.empty(d + 2 + (1 if est else 0))
        theta[0] = math.log(max(ye.var(), 1e-8))
        theta[1:1 + d] = math.log(0.2 * math.sqrt(d))
        if est:
            theta[1 + d] = math.log(1.5)
        fixed = self.nugget is not None
        theta[-1] = math.log(self.nugget if fixed else 0.1)
        theta = clip_theta(theta, d, self.lambda_max, est)
        self.theta0_ = theta.copy()
        state = {}

        def rebuild(th):
            Xs = Xe / np.exp(th[1:1 + d])
            order = maximin_order(Xs)
            nn = find_ordered_nn(Xs[order], m)
            state.update(order=order, nn=nn, groups=nn_groups(nn, m),
                         Xo=Xe[order], yo=ye[order], Zo=Ze[order])
            self.n_rebuilds += 1

        rebuild(theta)
        last = [np.exp(theta[1:1 + d]).copy()]
        lvt = math.log(max(ye.var(), 1e-8))

        def objective(th):
            th = clip_theta(np.array(th, dtype=float), d, self.lambda_max, est)
            if fixed:
                th[-1] = math.log(self.nugget)
            self.n_evals += 1
            ll, g, M, beta = vecchia_profile_loglik(
                th, state["Xo"], state["yo"], state["Zo"], state["groups"], self.nu,
                need_grad=True, var_penalty=self.var_penalty, log_var_target=lvt,
                estimate_nu=est)
            if fixed:
                g = g.copy(); g[-1] = 0.0
                M = M.copy(); M[-1, :] = 0.0; M[:, -1] = 0.0; M[-1, -1] = 1.0
            return ll, g, M, beta

        def reorder_hook(th, k):
            r = np.exp(clip_theta(np.array(th), d, self.lambda_max, est)[1:1 + d])
            if np.max(np.abs(np.log(r / last[0]))) < 1e-3:
                return False
            rebuild(th)
            last[0] = r.copy()
            return True

        theta, ll, beta = fisher_scoring(theta, objective, max_iter=self.max_iter,
                                         tol=self.tol, reorder_hook=reorder_hook,
                                         trace=self.trace)
        theta = clip_theta(np.array(theta), d, self.lambda_max, est)
        if fixed:
            theta[-1] = math.log(self.nugget)
        self.theta_, self.beta_, self.loglik_ = theta, beta, ll
        self._nu_fit = math.exp(theta[1 + d]) if est else self.nu     # gp.py:233
        self.state_ = state
        self.b_ = 1.0
        if self.var_correction and n >= 200:
            self.b_ = self._estimate_b(rng)
        return self

    def _predict_scaled(self, Xs_std, X_tr, resid, m, noise_free=True):
        """gp.py:242-277 with a brute-force sorted kNN."""
        ranges = self.ranges_
        var = math.exp(self.theta_[0])
        nug = math.exp(self.theta_[-1])
        m = min(m, X_tr.shape[0])
        nb = knn_query(X_tr / ranges, Xs_std / ranges, m)
        K = m + 1
        ns = Xs_std.shape[0]
        mean = np.empty(ns)
        varn = np.empty(ns)
        for a0, a1 in chunks(ns, K, 1):
            sl = slice(a0, a1)
            Xb = np.concatenate([X_tr[nb[sl]], Xs_std[sl][:, None, :]], axis=1)
            mask = np.ones((a1 - a0, K))
            if noise_free:
                mask[:, K - 1] = 0.0
            Sig, _ = block_cov(Xb, ranges, var, nug, self.nu_, derivs=False, nug_mask=mask)
            L, _ = batch_chol(Sig)
            E = np.zeros((a1 - a0, K, 1))
            E[:, K - 1, 0] = 1.0
            av = backsolve_LT(L, E)[:, :, 0]
            mean[sl] = -np.einsum("bk,bk->b", av[:, :m], resid[nb[sl]]) / av[:, K - 1]
            varn[sl] = 1.0 / av[:, K - 1] ** 2
        return mean, varn

    def predict(self, Xstar, return_std=False, return_var=False, m=None):
        """gp.py:279-304."""
        Xstar = np.ascontiguousarray(np.atleast_2d(Xstar), dtype=float)
        Xs = (Xstar - self._lo) / self._span
        m = self.m_pred if m is None else m
        Z_tr = design(self.X_, self.trend)
        resid = self.y_ - (Z_tr @ self.beta_ if self.beta_.size else 0.0)
        mu, va = self._predict_scaled(Xs, self.X_, resid, m)
        if self.beta_.size:
            mu = mu + design(Xs, self.trend) @ self.beta_
        mean = self._ymu + self._ysd * mu
        if not (return_std or return_var):
            return mean
        var = (self._ysd ** 2) * self.b_ * va
        if return_var:
            return mean, var
        return mean, np.sqrt(var)

    def _joint_factor(self, Xs, m):
        """gp.py:307-358."""
        n, d = self.X_.shape
        ns = Xs.shape[0]
        ranges = self.ranges_
        var = math.exp(self.theta_[0])
        nug = math.exp(self.theta_[-1])
        Xall = np.vstack([self.X_, Xs])
        Xsc = Xall / ranges
        order_obs = maximin_order(Xsc[:n])
        order_pred = maximin_order(Xsc[n:]) + n
        order = np.concatenate([order_obs, order_pred])
        Xo_sc = Xsc[order]
        nn = find_ordered_nn(Xo_sc, min(m, n), start=n)
        Xo = Xall[order]
        rows_i, cols_j, vals = [], [], []
        K = min(m, n) + 1
        pred_rows = np.arange(n, n + ns)
        for a0, a1 in chunks(ns, K, 1):
            r = pred_rows[a0:a1]
            block = np.concatenate([nn[r, 1:K], r[:, None]], axis=1)
            mask = np.ones((a1 - a0, K))
            mask[:, K - 1] = 0.0
            mask[:, :K - 1] = (block[:, :K - 1] < n).astype(float)
            Sig, _ = block_cov(Xo[block], ranges, var, nug, self.nu_, derivs=False, nug_mask=mask)
            L, _ = batch_chol(Sig)
            E = np.zeros((a1 - a0, K, 1))
            E[:, K - 1, 0] = 1.0
            av = backsolve_LT(L, E)[:, :, 0]
            rows_i.append(block.ravel())
            cols_j.append(np.repeat(np.arange(a0, a1), K))
            vals.append(av.ravel())
        Ufull = sparse.coo_matrix((np.concatenate(vals),
                                   (np.concatenate(rows_i), np.concatenate(cols_j))),
                                  shape=(n + ns, ns)).tocsc()
        return Ufull[:n, :], Ufull[n:, :], order[n:] - n, order_obs

    def predict_joint(self, Xstar, m=None, n_sim=0, exact_var=False, random_state=None):
        """gp.py:360-402."""
        Xstar = np.ascontiguousarray(np.atleast_2d(Xstar), dtype=float)
        Xs = (Xstar - self._lo) / self._span
        m = self.m_pred if m is None else m
        ns = Xs.shape[0]
        U_op, U_pp, perm, order_obs = self._joint_factor(Xs, m)
        inv = np.empty(ns, dtype=np.int64)
        inv[perm] = np.arange(ns)
        Z_tr = design(self.X_, self.trend)
        resid = self.y_ - (Z_tr @ self.beta_ if self.beta_.size else 0.0)
        Lt = U_pp.T.tocsr()
        rhs = -(U_op.T @ resid[order_obs])
        mu = spsolve_triangular(Lt, rhs, lower=True)[inv]
        if self.beta_.size:
            mu = mu + design(Xs, self.trend) @ self.beta_
        out = {"mean": self._ymu + self._ysd * mu}
        rng = np.random.default_rng(self.random_state if random_state is None else random_state)
        if exact_var:
            S = spsolve_triangular(Lt, np.eye(ns), lower=True)
            out["var"] = (self._ysd ** 2) * self.b_ * (S ** 2).sum(1)[inv]
        if n_sim > 0:
            eps = rng.standard_normal((ns, n_sim))
            dr = spsolve_triangular(Lt, eps, lower=True)[inv]
            samples = (mu[:, None] + math.sqrt(self.b_) * dr).T
            out["samples"] = self._ymu + self._ysd * samples
            if "var" not in out:
                out["var"] = out["samples"].var(0, ddof=1)
        return out

    def _estimate_b(self, rng):
        """gp.py:410-425."""
        n = self.X_.shape[0]
        n_in = min(2000, n // 5)
        idx = rng.permutation(n)
        te, tr = idx[:n_in], idx[n_in:]
        Z_tr = design(self.X_[tr], self.trend)
        resid = self.y_[tr] - (Z_tr @ self.beta_ if self.beta_.size else 0.0)
        mu, va = self._predict_scaled(self.X_[te], self.X_[tr], resid, min(self.m_pred, len(tr)))
        if self.beta_.size:
            mu = mu + design(self.X_[te], self.trend) @ self.beta_
        va = np.maximum(va, 1e-300)
        return max(float(np.mean((self.y_[te] - mu) ** 2 / va)), 1e-8)


# --------------------------------------------------------------------------
# dense closed forms used to pin the above
# --------------------------------------------------------------------------

def dense_cov(X, theta, nu):
    """Full covariance sigma^2 (M(r) + tau I) and its log-parameter derivatives."""
    d = X.shape[1]
    Sig, dS = block_cov(X[None], np.exp(theta[1:1 + d]), math.exp(theta[0]),
                        math.exp(theta[1 + d]), nu, derivs=True)
    return Sig[0], dS[0]


def exact_loglik(X, y, theta, nu, Z=None):
    """Exact GP loglik (profiled over beta by GLS when Z is given), its
    Fisher information 1/2 tr(K^-1 K_j K^-1 K_k) and beta."""
    K, dK = dense_cov(X, theta, nu)
    L = np.linalg.cholesky(K)
    if Z is not None and Z.shape[1] > 0:
        Zt = scipy.linalg.solve_triangular(L, Z, lower=True)
        yt = scipy.linalg.solve_triangular(L, y, lower=True)
        beta = np.linalg.solve(Zt.T @ Zt, Zt.T @ yt)
        r = y - Z @ beta
    else:
        beta = np.zeros(0)
        r = y
    a = scipy.linalg.solve_triangular(L, r, lower=True)
    ll = -0.5 * (len(y) * LOG2PI + 2 * np.log(np.diag(L)).sum() + a @ a)
    Ki = np.linalg.inv(K)
    A = [Ki @ dK[j] for j in range(dK.shape[0])]
    P = len(A)
    F = np.array([[0.5 * np.trace(A[j] @ A[k]) for k in range(P)] for j in range(P)])
    return ll, F, beta
