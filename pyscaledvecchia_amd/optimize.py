"""Fisher scoring with line-search fallback (host control loop).

Same rules as reference optimize.py:11-73: Fisher step ridged until it is
an ascent direction (<= 12 tries, ridge floor 1e-6, x10), normalised
gradient fallback, step norm capped at `max_step`, up to 20 halvings with
strict-increase acceptance (a LinAlgError halves the step), reorder hook at
k = 2, 4, 8, ..., stop when |alpha step . g| < tol or the search fails.
"""

from __future__ import annotations

import numpy as np

__all__ = ["_fisher_scoring"]


def _ascent_direction(g, M):
    diag = np.maximum(np.diag(M), 1e-8)
    ridge = 0.0
    for _ in range(12):
        try:
            cand = np.linalg.solve(M + ridge * np.diag(diag), g)
        except np.linalg.LinAlgError:
            ridge = max(1e-6, ridge * 10.0)
            continue
        if np.dot(cand, g) > 0 and np.all(np.isfinite(cand)):
            return cand
        ridge = max(1e-6, ridge * 10.0)
    return g / max(np.linalg.norm(g), 1e-12)


def _fisher_scoring(theta0, objective, max_iter=40, tol=1e-4, verbose=False,
                    reorder_hook=None, max_step=3.0, history=None):
    theta = np.array(theta0, dtype=float)
    ll, g, M, beta = objective(theta)
    next_reorder = 2
    for k in range(1, max_iter + 1):
        if reorder_hook is not None and k == next_reorder:
            next_reorder *= 2
            if reorder_hook(theta, k):
                ll, g, M, beta = objective(theta)
        step = _ascent_direction(g, M)
        nrm = np.linalg.norm(step)
        if nrm > max_step:
            step = step * (max_step / nrm)
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
        if verbose:
            print(f"    [FS {k:2d}] loglik = {ll:14.4f}   "
                  f"|step.grad| = {abs(np.dot(alpha * step, g)):.3e}")
        if history is not None:
            history.append((k, float(ll)))
        if not ok:
            break
        if abs(np.dot(alpha * step, g)) < tol:
            break
    return theta, ll, beta
