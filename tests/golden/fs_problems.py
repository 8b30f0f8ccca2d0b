"""Synthetic objectives that drive Fisher scoring (reference optimize.py:11-73)
through each of its branches.  Shared by the golden writer
(`write_fisher_scoring.py`) and the tests that replay the goldens.

Every problem is deterministic NumPy arithmetic.  An objective records the
argument of every call; a reorder hook records (theta, k) and its answer.

Branches covered (optimize.py line):
  * plain Fisher steps and the |alpha step . g| < tol stop (36, 70)
  * ridge ladder after a non-ascent solve (39-43)        -> "indefinite"
  * LinAlgError inside the ridged solve (37-38)           -> "singular"
  * every ridge fails -> normalised gradient (44-45)      -> "negdef"
  * non-finite candidate step (39)                        -> "nan_fisher"
  * step norm capped at max_step (47-49)                  -> "singular", "far"
  * LinAlgError inside the line search halves (56-58)     -> "raises"
  * non-finite loglik rejected (59)                       -> "nonfinite"
  * 20 halvings without increase ends the loop (68-69)    -> "stuck"
  * reorder hook at k = 2, 4, 8, ... re-evaluates (22-26) -> "reorder"
  * max_iter reached (19)                                 -> "slow"
"""

from __future__ import annotations

import numpy as np

__all__ = ["PROBLEMS", "make"]


class _Objective:
    """ll(theta) = -1/2 (theta - c)' A (theta - c) + w . sin(theta), its gradient,
    a (possibly perturbed) Fisher matrix and a 'beta' derived from theta."""

    def __init__(self, A, c, w, fisher=None, raise_radius=None, nan_radius=None,
                 wrong_grad_after=None):
        self.A = np.array(A, dtype=float)
        self.c = np.array(c, dtype=float)
        self.w = np.array(w, dtype=float)
        self.fisher = fisher
        self.raise_radius = raise_radius
        self.nan_radius = nan_radius
        self.wrong_grad_after = wrong_grad_after
        self.calls = []

    def __call__(self, theta):
        theta = np.array(theta, dtype=float)
        self.calls.append(theta.copy())
        r = theta - self.c
        if self.raise_radius is not None and np.linalg.norm(theta) > self.raise_radius:
            raise np.linalg.LinAlgError("synthetic Cholesky failure")
        ll = -0.5 * r @ self.A @ r + self.w @ np.sin(theta)
        if self.nan_radius is not None and np.linalg.norm(theta) > self.nan_radius:
            ll = np.nan
        g = -(self.A @ r) + self.w * np.cos(theta)
        if self.wrong_grad_after is not None and len(self.calls) > self.wrong_grad_after:
            g = -g
        M = self.A.copy() if self.fisher is None else self.fisher(theta, len(self.calls))
        beta = np.array([theta.sum(), theta @ theta])
        return ll, g, M, beta


class _Hook:
    def __init__(self, answers, obj, shift):
        self.answers = dict(answers)
        self.obj = obj
        self.shift = np.array(shift, dtype=float)
        self.calls = []

    def __call__(self, theta, k):
        ans = bool(self.answers.get(k, False))
        self.calls.append((np.array(theta, dtype=float).copy(), int(k), ans))
        if ans:   # the objective changes, as a rebuild changes the Vecchia structure
            self.obj.c = self.obj.c + self.shift
        return ans


def _spd(P, seed, cond=30.0):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((P, P)))
    ev = np.geomspace(1.0, cond, P)
    return (Q * ev) @ Q.T


def make(name):
    """Returns (theta0, objective, hook or None, kwargs of _fisher_scoring)."""
    P = 5
    A = _spd(P, 1)
    c = np.array([0.4, -0.3, 1.1, 0.2, -0.8])
    w = np.array([0.05, -0.02, 0.03, 0.01, 0.04])
    th0 = np.zeros(P)
    kw = dict(max_iter=40, tol=1e-4)
    hook = None
    if name == "plain":
        obj = _Objective(A, c, w)
    elif name == "indefinite":
        # early Fisher matrices are indefinite along the gradient direction u
        # (A - k u u' with k u'A^-1 u = 2 makes g'M^-1 g < 0), so the plain
        # solve is no ascent direction and the ridge ladder runs
        def fisher(theta, ncall):
            if ncall > 3:
                return A
            g = -(A @ (theta - c)) + w * np.cos(theta)
            u = g / np.linalg.norm(g)
            k = 1.5 / (u @ np.linalg.solve(A, u))
            return A - k * np.outer(u, u)
        obj = _Objective(A, c, w, fisher=fisher)
    elif name == "singular":
        # all-zero Fisher: solve() raises at ridge 0, ridge 1e-6 x diag floor
        # 1e-8 gives a huge step, which the cap shortens
        obj = _Objective(A, c, w, fisher=lambda th, n: np.zeros((P, P)) if n <= 2 else A)
    elif name == "negdef":
        # every ridged matrix stays negative definite -> normalised gradient
        obj = _Objective(A, c, w, fisher=lambda th, n: -100.0 * np.eye(P) if n <= 4 else A)
    elif name == "nan_fisher":
        def fisher(theta, ncall):
            if ncall == 1:
                M = A.copy()
                M[0, 1] = M[1, 0] = np.nan
                return M
            return A
        obj = _Objective(A, c, w, fisher=fisher)
    elif name == "raises":
        obj = _Objective(A, c * 4.0, w, fisher=lambda th, n: 0.05 * A, raise_radius=2.5)
    elif name == "nonfinite":
        obj = _Objective(A, c * 4.0, w, fisher=lambda th, n: 0.05 * A, nan_radius=2.5)
    elif name == "far":
        obj = _Objective(0.01 * A, c * 20.0, w)
    elif name == "stuck":
        # after a few calls the reported gradient points downhill, so no
        # halving raises the loglik and the loop ends
        obj = _Objective(A, c * 3.0, w, wrong_grad_after=1)
    elif name == "reorder":
        obj = _Objective(0.2 * A, c * 3.0, w, fisher=lambda th, n: 0.6 * A)
        hook = _Hook({2: True, 4: False, 8: True, 16: True}, obj, shift=[0.05, 0, -0.05, 0.02, 0])
        kw = dict(max_iter=40, tol=1e-9)
    elif name == "slow":
        obj = _Objective(A, c * 5.0, w, fisher=lambda th, n: 40.0 * A)
        kw = dict(max_iter=7, tol=0.0)
    else:
        raise KeyError(name)
    return th0, obj, hook, kw


PROBLEMS = ["plain", "indefinite", "singular", "negdef", "nan_fisher", "raises",
            "nonfinite", "far", "stuck", "reorder", "slow"]
