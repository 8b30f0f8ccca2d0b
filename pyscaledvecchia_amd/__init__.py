"""pyscaledvecchia_amd: the scaled-Vecchia hot path on AMD MI355X (gfx950).

Drop-in for the reference package `scaled_vecchia` (jdtuck/pyscaledvecchia):

    from pyscaledvecchia_amd import ScaledVecchiaGP
    gp = ScaledVecchiaGP(m_est=30, m_pred=140, nu=2.5).fit(X_train, y_train)
    mean, sd = gp.predict(X_test, return_std=True)
    draws    = gp.sample_joint(X_path, n_sim=100)

The maximin ordering, ordered nearest neighbours, every likelihood /
gradient / Fisher evaluation and the prediction blocks run in libsvh.so
(hand-written HIP for gfx950) through a ctypes C ABI (include/svh.h).
"""

from .gp import ScaledVecchiaGP
from .likelihood import VecchiaStructure, vecchia_profile_loglik
from .ordering import find_ordered_nn, maximin_order

__all__ = [
    "ScaledVecchiaGP",
    "maximin_order",
    "find_ordered_nn",
    "vecchia_profile_loglik",
    "VecchiaStructure",
]

__version__ = "0.1.0"
