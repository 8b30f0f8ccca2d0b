"""Profile Vecchia log-likelihood, gradient and Fisher information on the GPU.

Drop-in for scaled_vecchia.likelihood.vecchia_profile_loglik
(reference likelihood.py:66-175) for fixed nu in {0.5, 1.5, 2.5}.  The
structure (ordered inputs + neighbour groups) is uploaded once by
`VecchiaStructure`; each evaluation is one pass of the fused HIP kernels.
"""

from __future__ import annotations

import numpy as np

from . import _lib

__all__ = ["vecchia_profile_loglik", "VecchiaStructure", "DeviceRebuild"]

_CHUNK_BUDGET = 4_000_000     # _covariance.py:255


def _chunk_ids(groups, P):
    """Jitter-chunk id of every block, in group order: _chunks(len(rows), K, P)
    slicing of each group (_covariance.py:255-259, likelihood.py:111-114)."""
    ids = []
    base = 0
    for rows, nb in groups:
        K = nb.shape[1] + 1
        step = max(1, _CHUNK_BUDGET // max(1, K * K * max(P, 1)))
        local = np.arange(len(rows)) // step
        ids.append(base + local)
        base += int(local[-1]) + 1 if len(rows) else 0
    return (np.concatenate(ids) if ids else np.zeros(0, np.int64)), base


def _chunk_ids_rows(n, m, P):
    """_chunk_ids of the groups _nn_groups forms for the ordered neighbour
    lists of n rows with m neighbours (ordering.py:119-127): the groups of
    k = min(i, m) are rows 0..m-1 one by one, then rows m..n-1 together, so
    the chunk of row i is i below m and m + (i - m) // step above."""
    K = m + 1
    step = max(1, _CHUNK_BUDGET // max(1, K * K * max(P, 1)))
    i = np.arange(n)
    ids = np.where(i < m, i, m + (i - m) // step)
    return ids, (m + (n - m + step - 1) // step) if n > m else n


class VecchiaStructure:
    """Device-resident Vecchia structure: ordered inputs, responses, design
    and neighbour groups (the state gp.py:182-191 rebuilds)."""

    def __init__(self, Xo, yo, Zo, groups, ctx=None, estimate_nu=False):
        self.ctx = ctx or _lib.context()
        Xo = np.ascontiguousarray(Xo, dtype=np.float64)
        yo = np.ascontiguousarray(yo, dtype=np.float64)
        self.n, self.d = Xo.shape
        q = 0 if (Zo is None or Zo.shape[1] == 0) else Zo.shape[1]
        self.q = q
        groups = [(np.asarray(r), np.asarray(nb)) for r, nb in groups]
        mmax = max((nb.shape[1] for _, nb in groups), default=0)
        nblk = sum(len(r) for r, _ in groups)
        rows = np.empty(nblk, dtype=np.int64)
        nbr = np.full((nblk, max(mmax, 1)), -1, dtype=np.int64)
        kcnt = np.empty(nblk, dtype=np.int32)
        o = 0
        for r, nb in groups:
            B, k = len(r), nb.shape[1]
            rows[o:o + B] = r
            nbr[o:o + B, :k] = nb
            kcnt[o:o + B] = k
            o += B
        # jitter chunks of the gradient pass use P = d + 2, or d + 3 when the
        # smoothness is an extra parameter (likelihood.py:93, 114)
        P = self.d + (3 if estimate_nu else 2)
        cg, ncg = _chunk_ids(groups, P)
        cv, ncv = _chunk_ids(groups, 1)
        self.ctx.vecchia_setup(Xo, yo, None if q == 0 else Zo, rows, nbr, kcnt, max(mmax, 1),
                               cg, cv, max(ncg, 1), max(ncv, 1))
        self.nblk = nblk
        self.mmax = mmax

    def __call__(self, theta, nu, need_grad=True, var_penalty=0.0, log_var_target=0.0,
                 precision="fp64"):
        return self.ctx.vecchia_eval(theta, nu, need_grad=need_grad, var_penalty=var_penalty,
                                     log_var_target=log_var_target, precision=precision)


class DeviceRebuild:
    """The rebuilds of one fit's Vecchia structure on the device
    (svh_vecchia_rebuild): the estimation data are uploaded by the first call,
    every call reorders them at new ranges (maximin order, ordered neighbours,
    structure) without moving the neighbour lists through the host.  Calling
    the object's `structure` evaluates like a VecchiaStructure."""

    def __init__(self, Xe, ye, Ze, m, ctx=None):
        self.ctx = ctx or _lib.context()
        self.Xe = np.ascontiguousarray(Xe, dtype=np.float64)
        self.ye = np.ascontiguousarray(ye, dtype=np.float64)
        self.Ze = None if (Ze is None or Ze.shape[1] == 0) else np.ascontiguousarray(Ze, float)
        self.m = int(m)

    @staticmethod
    def supported(d, m):
        return d <= 12 and m + 1 <= 127

    def __call__(self, lam):
        """Reorder at ranges `lam`; returns the maximin order.  The first
        point (ordering.py:72, the row closest to the centroid of X / lam) is
        found on the device with NumPy's arithmetic (first = -1)."""
        lam = np.ascontiguousarray(lam, dtype=np.float64)
        first = -1
        fresh = getattr(self.ctx, "_rb_owner", None) is not self
        order = self.ctx.vecchia_rebuild(self.Xe if fresh else None, self.ye, self.Ze, lam,
                                         first, self.m, upload=fresh)
        self.ctx._rb_owner = self
        return order

    def structure(self, theta, nu, need_grad=True, var_penalty=0.0, log_var_target=0.0,
                  precision="fp64"):
        return self.ctx.vecchia_eval(theta, nu, need_grad=need_grad, var_penalty=var_penalty,
                                     log_var_target=log_var_target, precision=precision)


def _check_nu(nu):
    if not (nu is not None and np.isfinite(nu) and nu > 0):
        raise ValueError("nu must be positive (closed forms at 0.5, 1.5, 2.5, the Bessel "
                         "form otherwise)")


def vecchia_profile_loglik(theta, Xo, yo, Zo, groups, nu, estimate_nu=False,
                           need_grad=True, var_penalty=0.0, log_var_target=0.0,
                           precision="fp64"):
    """Profile Vecchia loglikelihood, gradient and Fisher information
    (likelihood.py:66-175).  theta = [log sigma^2, log lambda_1..d, log tau]
    (P = d + 2), or [log sigma^2, log lambda_1..d, log nu, log tau] (P = d + 3)
    with estimate_nu, where `nu` is then ignored.  Returns (ll, grad, fisher,
    beta), or (ll, None, None, beta) when need_grad is False."""
    if not estimate_nu:
        _check_nu(nu)
    st = VecchiaStructure(Xo, yo, Zo, groups, estimate_nu=estimate_nu)
    return st(theta, 0.0 if estimate_nu else nu, need_grad=need_grad,
              var_penalty=var_penalty, log_var_target=log_var_target, precision=precision)
