"""Fisher scoring (reference optimize.py:11-73) replayed against the
reference's own trajectories.

tests/golden/fisher_scoring.npz holds, for each synthetic objective of
tests/golden/fs_problems.py, every argument the reference's _fisher_scoring
passed to the objective, every reorder-hook call, and its result
(written by tests/golden/write_fisher_scoring.py from the reference source).
The oracle's restatement and the product's host loop must reproduce them call
for call.  The device state machine of the batched fitter is held to the host
loop call for call in tests/test_gpu_batch.py.
"""

import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import fs_problems  # noqa: E402

from oracle import vecchia_np as O  # noqa: E402
from pyscaledvecchia_amd.optimize import _fisher_scoring  # noqa: E402

GOLD = np.load(os.path.join(HERE, "golden", "fisher_scoring.npz"))


def _oracle(th0, obj, hook, kw):
    return O.fisher_scoring(th0, obj, reorder_hook=hook, **kw)


def _product(th0, obj, hook, kw):
    return _fisher_scoring(th0, obj, reorder_hook=hook, **kw)


@pytest.mark.parametrize("impl", [_oracle, _product], ids=["oracle", "product"])
@pytest.mark.parametrize("name", fs_problems.PROBLEMS)
def test_trajectory_matches_reference(name, impl):
    th0, obj, hook, kw = fs_problems.make(name)
    theta, ll, beta = impl(th0, obj, hook, kw)
    calls = np.array(obj.calls)
    ref = GOLD[f"{name}/calls"]
    assert calls.shape == ref.shape, f"{name}: {len(calls)} objective calls, reference {len(ref)}"
    # identical NumPy/LAPACK arithmetic: equal to the last bit on one machine;
    # 1e-14 leaves room for a LAPACK build that orders its sums differently
    np.testing.assert_allclose(calls, ref, rtol=1e-14, atol=1e-14)
    np.testing.assert_allclose(theta, GOLD[f"{name}/theta"], rtol=1e-14, atol=1e-14)
    np.testing.assert_allclose(beta, GOLD[f"{name}/beta"], rtol=1e-14, atol=1e-14)
    ref_ll = float(GOLD[f"{name}/ll"])
    assert (np.isnan(ll) and np.isnan(ref_ll)) or ll == pytest.approx(ref_ll, rel=1e-14, abs=1e-14)
    if hook is not None:
        np.testing.assert_array_equal([k for _, k, _ in hook.calls], GOLD[f"{name}/hook_k"])
        np.testing.assert_array_equal([a for _, _, a in hook.calls], GOLD[f"{name}/hook_ans"])
        np.testing.assert_allclose([t for t, _, _ in hook.calls], GOLD[f"{name}/hook_theta"],
                                   rtol=1e-14, atol=1e-14)


def test_goldens_exercise_each_branch():
    """The recorded trajectories reach the branches the problems are named for."""
    calls = {n: GOLD[f"{n}/calls"] for n in fs_problems.PROBLEMS}
    # raises / nonfinite: candidates outside the radius were proposed and halved
    for n in ("raises", "nonfinite"):
        norms = np.linalg.norm(calls[n], axis=1)
        assert (norms > 2.5).sum() >= 5 and (norms <= 2.5).sum() >= 5
    # stuck: the loop ends after 20 rejected halvings
    th = calls["stuck"]
    assert len(th) >= 21
    d = th[-20:] - th[-21]
    assert np.allclose(d / np.linalg.norm(d, axis=1)[:, None], d[0] / np.linalg.norm(d[0]))
    # reorder: hook called at k = 2, 4, 8, 16 with the scripted answers
    np.testing.assert_array_equal(GOLD["reorder/hook_k"][:4], [2, 4, 8, 16])
    # slow: max_iter = 7 iterations of one accepted full step each
    assert len(calls["slow"]) == 8
    # far / singular: capped steps of norm 3 (max_step)
    steps = np.linalg.norm(np.diff(calls["far"], axis=0), axis=1)
    assert np.isclose(steps.max(), 3.0)
    steps = np.linalg.norm(calls["singular"][1] - calls["singular"][0])
    assert np.isclose(steps, 3.0) or np.isclose(steps, 1.5)
    # negdef: the first step is a normalised gradient (length 1 before halvings)
    s = np.linalg.norm(calls["negdef"][1] - calls["negdef"][0])
    assert np.isclose(s, 1.0) or np.isclose(s * 2 ** np.round(np.log2(1 / s)), 1.0)
