"""ScaledVecchiaGP: drop-in for scaled_vecchia.gp.ScaledVecchiaGP.

Same fields, defaults, fitted attributes, methods, exceptions and numpy RNG
call order as the reference (gp.py:40-444).  The host keeps the control
flow (standardisation, subsampling, Fisher scoring, reorder schedule,
variance correction); the hot path -- maximin order, ordered neighbours,
every likelihood evaluation, the prediction blocks and their kNN query --
runs on the GPU through libsvh.so.

Additive knobs (defaults reproduce the reference): `precision` ("fp64" or
"mixed"), `theta0` (a warm start for Fisher scoring), `warm_start` (a learned
starting point: a WarmStart network, or "learned" for the packaged one),
`device`.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .likelihood import DeviceRebuild, VecchiaStructure
from .optimize import _fisher_scoring
from .ordering import _nn_groups, find_ordered_nn, maximin_order

__all__ = ["ScaledVecchiaGP"]

_CHUNK_BUDGET = 4_000_000


@dataclass
class ScaledVecchiaGP:
    """Scaled Vecchia GP emulator (Katzfuss, Guinness & Lawrence, 2020).

    m_est / m_pred : conditioning-set sizes for estimation / prediction
    n_est          : subsample size used for estimation
    nu             : Matern smoothness: any positive value (closed forms at
                     0.5, 1.5, 2.5, the Bessel form otherwise), or None to
                     estimate it as an extra parameter
    trend          : 'zero', 'constant' or 'linear'
    nugget         : None -> estimated relative nugget; float -> fixed
    var_correction : estimate the predictive-variance factor b (Sec. 3.4)
    lambda_max     : ranges above this are treated as infinite
    precision      : 'fp64' (default), 'mixed' or 'mixed_f32cov': during the
                     search the derivative terms dS a, the W solves and the
                     Fisher / cross products run in FP32 ('mixed_f32cov' also
                     assembles the off-diagonal covariance entries in FP32);
                     Cholesky, log-determinant, GLS stay FP64; then FP64
                     refinement steps; predictions are FP64 either way
    theta0         : optional initial log-parameters for Fisher scoring
    warm_start     : None (default start), "learned" (the packaged network) or
                     a WarmStart instance: the data are also fitted from the
                     initial log-parameters the network predicts from summary
                     features, and that fit is kept only if it ends at a
                     higher log-likelihood than the reference start's fit
    """

    m_est: int = 30
    m_pred: int = 140
    n_est: int = 5000
    nu: float | None = 2.5
    trend: str = "constant"
    nugget: float | None = None
    var_penalty: float = 1.0
    var_correction: bool = True
    lambda_max: float = 1e3
    max_iter: int = 40
    tol: float = 1e-4
    random_state: int | None = 0
    verbose: bool = False
    precision: str = "fp64"
    theta0: np.ndarray | None = field(default=None, repr=False)
    warm_start: object = field(default=None, repr=False)
    device: int = 0

    X_: np.ndarray = field(default=None, init=False, repr=False)
    y_: np.ndarray = field(default=None, init=False, repr=False)
    theta_: np.ndarray = field(default=None, init=False, repr=False)
    beta_: np.ndarray = field(default=None, init=False, repr=False)
    b_: float = field(default=1.0, init=False)
    d_: int = field(default=None, init=False, repr=False)
    _estimate_nu: bool = field(default=False, init=False, repr=False)
    _nu_fit: float = field(default=None, init=False, repr=False)

    # ---------------- helpers (gp.py:94-135) ------------------------------
    def _design(self, X):
        if self.trend == "zero":
            return np.zeros((X.shape[0], 0))
        if self.trend == "constant":
            return np.ones((X.shape[0], 1))
        if self.trend == "linear":
            return np.column_stack([np.ones(X.shape[0]), X])
        raise ValueError("trend must be 'zero', 'constant' or 'linear'")

    @property
    def variance_(self):
        return math.exp(self.theta_[0])

    @property
    def ranges_(self):
        return np.exp(self.theta_[1:1 + self.d_])

    @property
    def relevance_(self):
        """1 / lambda_l -- how strongly input l affects the response."""
        return 1.0 / self.ranges_

    @property
    def nu_(self):
        return self._nu_fit

    @property
    def nugget_(self):
        return math.exp(self.theta_[-1])

    def _clip_theta(self, theta, d, estimate_nu=False):
        hi = math.log(self.lambda_max) if np.isfinite(self.lambda_max) else np.inf
        theta[1:1 + d] = np.clip(theta[1:1 + d], math.log(1e-8), hi)
        if estimate_nu:
            # smoothness kept where the Bessel evaluations are well conditioned
            theta[1 + d] = np.clip(theta[1 + d], math.log(0.3), math.log(10.0))
        theta[-1] = np.clip(theta[-1], math.log(1e-12), math.log(1e3))
        return theta

    def _ctx(self):
        return _lib.context(self.device)

    def _check_nu(self):
        if self.nu is not None and not (np.isfinite(self.nu) and self.nu > 0):
            raise ValueError("nu must be positive, or None to estimate it")

    # ---------------- fitting (gp.py:138-239) -----------------------------
    def _prepare(self, X, y):
        """Validation, standardisation, subsample and initial parameters
        (gp.py:139-177).  Returns the estimation data and the numpy RNG in the
        state the reference leaves it (later used by the b-step)."""
        X = np.ascontiguousarray(np.atleast_2d(X), dtype=float)
        y = np.asarray(y, dtype=float).ravel()
        n, d = X.shape
        if y.shape[0] != n:
            raise ValueError("X and y have incompatible shapes")
        self._design(X[:1])                    # validates `trend` early
        self._check_nu()
        est = self.nu is None                  # gp.py:145
        self._estimate_nu = est
        if est and self.warm_start is not None and self.theta0 is None:
            raise ValueError("warm_start needs a fixed nu")
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
        Ze = self._design(Xe)
        m = min(self.m_est, len(sub) - 1)

        fixed_nug = self.nugget is not None
        P = d + 2 + (1 if est else 0)
        default = np.empty(P)
        default[0] = math.log(max(ye.var(), 1e-8))
        default[1:1 + d] = math.log(0.2 * math.sqrt(d))
        if est:
            default[1 + d] = math.log(1.5)     # gp.py:174-175
        default[-1] = math.log(self.nugget if fixed_nug else 0.1)
        default = self._clip_theta(default, d, est)
        if self.theta0 is not None:
            theta = np.array(self.theta0, dtype=float).copy()
            if theta.shape != (P,):
                raise ValueError(f"theta0 must have {P} entries")
        elif self.warm_start is not None:
            theta = getattr(self, "_ws_theta0", None)
            if theta is None:
                theta = self._warm_start_net().theta0(Xe, ye, self.nu)
            self._ws_theta0 = None
        else:
            theta = default.copy()
        if fixed_nug:
            theta[-1] = math.log(self.nugget)
        theta = self._clip_theta(theta, d, est)
        return dict(n=n, d=d, rng=rng, Xe=Xe, ye=ye, Ze=Ze, m=m, theta=theta,
                    theta_default=default, fixed_nug=fixed_nug,
                    lvt=math.log(max(ye.var(), 1e-8)))

    def _warm_start_net(self):
        from .warmstart import WarmStart
        if isinstance(self.warm_start, str):
            if self.warm_start != "learned":
                raise ValueError("warm_start must be None, 'learned' or a WarmStart")
            self.warm_start = WarmStart.load(device=self.device)
        return self.warm_start

    def _order_structure(self, Xe, theta, m):
        """Maximin order and ordered neighbours in the scaled space
        (gp.py:182-191), on the GPU."""
        d = Xe.shape[1]
        Xs = Xe / np.exp(theta[1:1 + d])
        order = maximin_order(Xs)
        nn = find_ordered_nn(Xs[order], m)
        return order, nn

    def _finish(self, theta, ll, beta, rng, n):
        """Final clip, fitted attributes and the b-step (gp.py:227-239)."""
        d = self.d_
        est = self._estimate_nu
        theta = self._clip_theta(np.array(theta), d, est)
        if self.nugget is not None:
            theta[-1] = math.log(self.nugget)
        self.theta_ = theta
        self.beta_ = beta
        self.loglik_ = ll
        self._nu_fit = math.exp(theta[1 + d]) if est else self.nu     # gp.py:233
        self.b_ = 1.0
        if self.var_correction and n >= 200:
            self.b_ = self._estimate_b(rng)
        return self

    def fit(self, X, y):
        if self.theta0 is None and self.warm_start is not None:
            return self._fit_learned(X, y)
        prep = self._prepare(X, y)
        n, d, rng = prep["n"], prep["d"], prep["rng"]
        Xe, ye, Ze, m, theta = prep["Xe"], prep["ye"], prep["Ze"], prep["m"], prep["theta"]
        fixed_nug = prep["fixed_nug"]

        ctx = self._ctx()
        state = {}
        self.n_rebuilds_ = 0
        self.n_evals_ = 0

        est = self._estimate_nu
        # the smoothness argument of the evaluations: 0 when it is estimated
        # (theta[1 + d] = log nu)
        nu_arg = 0.0 if est else self.nu
        dev = DeviceRebuild(Xe, ye, Ze, m, ctx=ctx) if DeviceRebuild.supported(d, m) else None

        def rebuild(th):
            if dev is not None:
                # maximin order, ordered neighbours and structure on the device
                state["order"] = dev(np.exp(np.asarray(th, float)[1:1 + d]))
                state["struct"] = dev.structure
            else:
                order, nn = self._order_structure(Xe, th, m)
                groups = _nn_groups(nn, m)
                state["order"] = order
                state["struct"] = VecchiaStructure(Xe[order], ye[order], Ze[order], groups,
                                                   ctx=ctx, estimate_nu=est)
            self.n_rebuilds_ += 1

        rebuild(theta)
        last_ranges = np.exp(theta[1:1 + d]).copy()
        log_var_target = prep["lvt"]

        def objective(th, precision=self.precision):
            th = self._clip_theta(np.array(th, dtype=float), d, est)
            if fixed_nug:
                th[-1] = math.log(self.nugget)
            self.n_evals_ += 1
            ll, g, M, beta = state["struct"](th, nu_arg, need_grad=True,
                                             var_penalty=self.var_penalty,
                                             log_var_target=log_var_target,
                                             precision=precision)
            if fixed_nug:
                g = g.copy(); g[-1] = 0.0
                M = M.copy(); M[-1, :] = 0.0; M[:, -1] = 0.0; M[-1, -1] = 1.0
            return ll, g, M, beta

        def reorder_hook(th, k):
            nonlocal last_ranges
            r = np.exp(self._clip_theta(np.array(th), d, est)[1:1 + d])
            if np.max(np.abs(np.log(r / last_ranges))) < 1e-3:
                return False
            if self.verbose:
                print(f"    [reorder at iter {k}] ranges = " + np.array2string(r, precision=3))
            rebuild(th)
            last_ranges = r.copy()
            return True

        self.history_ = []
        self.warm_start_used_ = False
        theta, ll, beta = _fisher_scoring(theta, objective, max_iter=self.max_iter,
                                          tol=self.tol, verbose=self.verbose,
                                          reorder_hook=reorder_hook, history=self.history_)
        self.n_iter_ = len(self.history_)
        self.n_refine_ = 0
        if self.precision not in ("fp64", "mixed", "mixed_f32cov"):
            raise ValueError("precision must be 'fp64', 'mixed' or 'mixed_f32cov'")
        if self.precision != "fp64":
            # Refinement: the mixed search used FP64 log-likelihoods but FP32
            # derivative products, so its stationary point is that of a
            # slightly perturbed gradient.  Fisher scoring with FP64
            # derivatives on the final structure moves it to the FP64 one
            # (one or two steps from this start).
            n0 = len(self.history_)
            theta, ll, beta = _fisher_scoring(
                theta, lambda th: objective(th, "fp64"), max_iter=self.max_iter,
                tol=self.tol, verbose=self.verbose, reorder_hook=None, history=self.history_)
            self.n_refine_ = len(self.history_) - n0
        self._fit_state = state
        return self._finish(theta, ll, beta, rng, n)

    def _fit_learned(self, X, y):
        """Fit with the learned starting point, judged by where the fit ends:
        the data are fitted from the learned start and from the reference
        start (gp.py:169-177) and the fit with the higher final
        log-likelihood is kept (the reference start on ties), so a learned
        start never returns a worse fit than the reference would.  Both fits
        share the batched fitter's launches when its limits allow (d <= 10,
        trend columns <= 11, m_est <= 31); otherwise they run one after the
        other."""
        from dataclasses import replace
        d = np.atleast_2d(X).shape[1]
        if (d <= 10 and self._design(np.zeros((1, d))).shape[1] <= 11 and self.m_est <= 31
                and self.nu in (0.5, 1.5, 2.5)):
            from .batch import BatchedFit
            BatchedFit([self], device=self.device).fit([X], [y])
            return self
        prep = self._prepare(X, y)          # learned start from the network
        try:
            learned = replace(self, theta0=prep["theta"], warm_start=None).fit(X, y)
        except np.linalg.LinAlgError:
            learned = None          # the learned start's blocks fail: reference start
        default = replace(self, theta0=None, warm_start=None).fit(X, y)
        use = learned if learned is not None and learned.loglik_ > default.loglik_ else default
        for k, v in vars(use).items():
            if k not in ("theta0", "warm_start"):
                setattr(self, k, v)
        self.warm_start_used_ = use is learned
        self.n_evals_both_ = default.n_evals_ + (learned.n_evals_ if learned is not None else 0)
        return self

    # ---------------- prediction (gp.py:242-304) ----------------------------
    def _predict_scaled(self, Xs_std, X_tr, y_tr_resid, m, noise_free=True):
        """Marginal Vecchia predictions from the m nearest training points
        (gp.py:242-277): kNN query and K = m+1 blocks on the GPU."""
        m = min(m, X_tr.shape[0])
        # FP64 blocks whatever the fit's precision: with tiny nuggets the
        # K = m + 1 blocks are ill-conditioned and FP32 trailing updates move
        # the predictive variance (and the b correction) visibly
        return self._ctx().predict(X_tr, y_tr_resid, Xs_std, m, self._theta_cov(), self.nu_,
                                   noise_free=noise_free, precision="fp64")

    def predict(self, Xstar, return_std=False, return_var=False, m=None):
        self._check_fitted()
        Xstar = np.ascontiguousarray(np.atleast_2d(Xstar), dtype=float)
        Xs = (Xstar - self._lo) / self._span
        m = self.m_pred if m is None else m
        Z_tr = self._design(self.X_)
        resid = self.y_ - (Z_tr @ self.beta_ if self.beta_.size else 0.0)
        mu, va = self._predict_scaled(Xs, self.X_, resid, m)
        if self.beta_.size:
            mu = mu + self._design(Xs) @ self.beta_
        mean = self._ymu + self._ysd * mu
        if not (return_std or return_var):
            return mean
        var = (self._ysd ** 2) * self.b_ * va
        if return_var:
            return mean, var
        return mean, np.sqrt(var)

    # ---------------- joint prediction (gp.py:307-407) ----------------------
    def _joint_factor(self, Xs, m):
        """Sparse inverse-Cholesky columns of the prediction block with
        observations ordered first (gp.py:307-358)."""
        n, d = self.X_.shape
        ns = Xs.shape[0]
        ranges = self.ranges_
        Xall = np.vstack([self.X_, Xs])
        Xsc = Xall / ranges
        # The reference orders the observations by maximin too (gp.py:321), an
        # O(n^2 d) pass that changes no output: every prediction point's
        # neighbours are chosen among all observations whatever their order
        # (sets by (distance, position): only exact distance ties could pick
        # another neighbour), and the order only permutes U_op's rows together
        # with resid[order_obs].  The observations keep their input order.
        order_obs = np.arange(n)
        order_pred = maximin_order(Xsc[n:]) + n
        order = np.concatenate([order_obs, order_pred])
        Xo_sc = Xsc[order]
        mm = min(m, n)
        # rows n.. of find_ordered_nn(Xo_sc, mm, start=n) (gp.py:326): the only
        # rows read below
        nn = self._ctx().ordered_knn_rows(Xo_sc, mm, n)
        Xo = Xall[order]
        K = mm + 1
        pred_rows = np.arange(n, n + ns)
        block = np.concatenate([nn[:, 1:K], pred_rows[:, None]], axis=1)
        mask = np.ones((ns, K))
        mask[:, K - 1] = 0.0
        mask[:, :K - 1] = (block[:, :K - 1] < n).astype(float)
        av = self._ctx().block_columns(Xo, block, mask, self._theta_cov(), self.nu_,
                                       precision="fp64")
        # U's column i holds av[i] at rows block[i]; the solves with
        # Lt = U_pp^T run on the device (svh_joint_solve)
        return block, av, order[n:] - n, order_obs

    def predict_joint(self, Xstar, m=None, n_sim=0, exact_var=False, random_state=None):
        """Joint Vecchia predictive distribution (gp.py:360-402)."""
        self._check_fitted()
        Xstar = np.ascontiguousarray(np.atleast_2d(Xstar), dtype=float)
        Xs = (Xstar - self._lo) / self._span
        m = self.m_pred if m is None else m
        ns = Xs.shape[0]
        block, av, perm, order_obs = self._joint_factor(Xs, m)
        n = self.X_.shape[0]
        inv = np.empty(ns, dtype=np.int64)
        inv[perm] = np.arange(ns)
        Z_tr = self._design(self.X_)
        resid = self.y_ - (Z_tr @ self.beta_ if self.beta_.size else 0.0)
        ctx = self._ctx()
        # a fresh generator whose only draw is eps (gp.py:389-396): drawn
        # first, so the mean, the paths and the exact variances are one solve
        rng = np.random.default_rng(self.random_state if random_state is None else random_state)
        eps = rng.standard_normal((ns, n_sim)) if n_sim > 0 else None
        mu, dr, var = ctx.joint_solve(n, block, av, resid=resid[order_obs], rhs=eps,
                                      exact_var=exact_var)
        mu = mu[inv]
        if self.beta_.size:
            mu = mu + self._design(Xs) @ self.beta_
        out = {"mean": self._ymu + self._ysd * mu}
        if exact_var:
            out["var"] = (self._ysd ** 2) * self.b_ * var[inv]
        if n_sim > 0:
            dr = dr[inv]
            samples = (mu[:, None] + math.sqrt(self.b_) * dr).T
            out["samples"] = self._ymu + self._ysd * samples
            if "var" not in out:
                out["var"] = out["samples"].var(0, ddof=1)
        return out

    def sample_joint(self, Xstar, n_sim=100, m=None, random_state=None):
        """Draw joint sample paths from the predictive distribution."""
        return self.predict_joint(Xstar, m=m, n_sim=n_sim, random_state=random_state)["samples"]

    # ---------------- variance correction (gp.py:410-425) -------------------
    def _estimate_b(self, rng):
        n = self.X_.shape[0]
        n_in = min(2000, n // 5)
        idx = rng.permutation(n)
        te, tr = idx[:n_in], idx[n_in:]
        Z_tr = self._design(self.X_[tr])
        resid = self.y_[tr] - (Z_tr @ self.beta_ if self.beta_.size else 0.0)
        mu, va = self._predict_scaled(self.X_[te], self.X_[tr], resid,
                                      min(self.m_pred, len(tr)))
        if self.beta_.size:
            mu = mu + self._design(self.X_[te]) @ self.beta_
        va = np.maximum(va, 1e-300)
        b = float(np.mean((self.y_[te] - mu) ** 2 / va))
        return max(b, 1e-8)

    def _theta_cov(self):
        """[log sigma^2, log lambda, log tau]: the parameters of the prediction
        blocks (the fitted smoothness goes separately, as nu_)."""
        th = self.theta_
        if self._estimate_nu:
            th = np.concatenate([th[:1 + self.d_], th[-1:]])
        return th

    def _check_fitted(self):
        if self.theta_ is None:
            raise RuntimeError("call .fit(X, y) first")

    def summary(self):
        self._check_fitted()
        nu_label = "estimated" if self._estimate_nu else "fixed"
        s = [f"ScaledVecchiaGP(nu={'estimate' if self.nu is None else self.nu}, "
             f"m_est={self.m_est}, m_pred={self.m_pred}, "
             f"trend='{self.trend}')",
             f"  Vecchia loglik (standardised y) : {self.loglik_:.3f}",
             f"  sigma^2 : {self.variance_:.6g}",
             f"  nu      : {self.nu_:.4g}  ({nu_label})",
             f"  nugget  : {self.nugget_:.6g}  (relative)",
             f"  b       : {self.b_:.4g}  (variance correction)",
             "  input relevances 1/lambda_l (unit-box scaling):"]
        for l, rel in enumerate(self.relevance_):
            s.append(f"      x[{l}] : {rel:10.4f}")
        return "\n".join(s)
