"""CPU oracle for the scaled-Vecchia hot path.  TEST INFRASTRUCTURE ONLY.

This package is the checker, never the product: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it.  The product package ``pyscaledvecchia_amd`` never imports it and
has no CPU fallback.

``vecchia_np`` is a plain NumPy FP64 restatement of the reference's algorithm
(jdtuck/pyscaledvecchia, ``src/scaled_vecchia/*.py``); every function cites the
reference file:line it follows.  ``liboracle_c`` (built from ``csrc/oracle.c``
with ``make -C oracle``) restates the O(n^2) ordering/neighbour loops and the
per-block likelihood in C for speed; it is checked against ``vecchia_np``.

How each oracle function is pinned (checked against something other than a
retyping of its own formula) is listed in ``PINS`` below; the tests in
``tests/test_oracle_*.py`` carry out each pin.
"""

PINS = {
    "matern_corr / matern_g": "scipy.special.kv general Matern at nu=.5/1.5/2.5; "
                              "central finite difference of M for g=M'/r",
    "maximin_order": "pure-Python brute force of the definition (tiny n); "
                     "permutation + determinism invariants (reference tests)",
    "maximin_first": "known answers: middle point of an evenly spaced line (lower index "
                     "of the two middle ones for even n), a row placed at the centroid",
    "find_ordered_nn": "pure-Python brute force per row (reference test_ordering); "
                       "structure: self column, -1 padding, start= invariance",
    "block_cov": "central finite differences of Sigma(theta) in log-parameters; "
                 "closed form at r=0",
    "batch_chol / solves": "numpy.linalg.cholesky, scipy.linalg.solve_triangular; "
                           "jitter ladder forced by duplicate rows",
    "vecchia_profile_loglik": "m=n-1 equals the dense exact GP loglik; FD gradient; "
                              "Fisher symmetric/PSD and equal to the dense "
                              "1/2 tr(K^-1 K_j K^-1 K_k) at m=n-1; beta equals dense GLS",
    "fisher_scoring": "the reference's own trajectories: every objective argument, "
                      "reorder-hook call and result of optimize.py:11-73 on eleven "
                      "synthetic objectives that reach each branch (ridge ladder, "
                      "LinAlgError in the solve and in the line search, gradient "
                      "fallback, non-finite step and loglik, step cap, failed search, "
                      "reorder, max_iter), tests/golden/fisher_scoring.npz",
    "predict_scaled": "dense conditional Gaussian on the m nearest neighbours",
    "predict_joint": "exact dense GP conditional for one point with m >= n; for "
                     "several points the dense Vecchia joint law written from its "
                     "definition (conditional regressions on brute-force neighbour "
                     "sets, noise-free prediction neighbours): mean, variance and "
                     "paths from the same normal draws",
    "estimate_b": "numerical minimiser of the held-out score on a split rebuilt by "
                  "hand from the RNG call order, dense marginal predictions; floor "
                  "and n < 200 cases",
    "mlp adam_step": "torch.optim.AdamW over six steps with weight decay",
}
