"""Writes fisher_scoring.npz: call-by-call trajectories of the reference's
Fisher scoring (src/scaled_vecchia/optimize.py:11-73) on the synthetic
objectives of fs_problems.py.

Run in the build container only, where the reference is mounted:

    python tests/golden/write_fisher_scoring.py [/root/reference]

optimize.py imports only numpy, so it is loaded by file path (no numba, no
package import).  The reference itself never travels with the repository;
only the vectors it produced do.
"""

from __future__ import annotations

import importlib.util
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.dont_write_bytecode = True      # the problems module is always read from source
sys.path.insert(0, HERE)
import fs_problems  # noqa: E402


def _load_reference(root):
    path = os.path.join(root, "src", "scaled_vecchia", "optimize.py")
    spec = importlib.util.spec_from_file_location("_ref_optimize", path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod._fisher_scoring


def main():
    root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    fs = _load_reference(root)
    out = {}
    for name in fs_problems.PROBLEMS:
        th0, obj, hook, kw = fs_problems.make(name)
        theta, ll, beta = fs(th0, obj, reorder_hook=hook, **kw)
        out[f"{name}/calls"] = np.array(obj.calls)
        out[f"{name}/theta"] = np.asarray(theta, dtype=float)
        out[f"{name}/ll"] = np.array(ll, dtype=float)
        out[f"{name}/beta"] = np.asarray(beta, dtype=float)
        if hook is not None:
            out[f"{name}/hook_theta"] = np.array([t for t, _, _ in hook.calls])
            out[f"{name}/hook_k"] = np.array([k for _, k, _ in hook.calls], dtype=np.int64)
            out[f"{name}/hook_ans"] = np.array([a for _, _, a in hook.calls], dtype=bool)
        print(f"{name:11s} calls={len(obj.calls):3d} ll={float(ll):.12g}")
    np.savez(os.path.join(HERE, "fisher_scoring.npz"), **out)


if __name__ == "__main__":
    main()
