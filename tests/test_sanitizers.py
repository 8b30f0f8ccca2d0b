"""CPU builds under AddressSanitizer and UndefinedBehaviorSanitizer
(tests/sanitize/Makefile), run on small inputs:

* the C oracle (oracle/csrc/oracle.c) on edge cases (n = 1, 2; m = 0; m >= n;
  duplicated points; d = 12; several thread counts);
* the library's host code, compiled host-only (no kernels), called through
  the C ABI: every entry point of include/svh.h that takes a context refuses a
  null one, context creation refuses bad devices, and the host Matern export
  checks its arguments and matches the closed forms.

No GPU is used.  Skipped where the sanitizer sources or the compilers are
absent.
"""

import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = os.path.join(ROOT, "tests", "sanitize")


def _context_entry_points():
    """(name, [parameter types]) of every `int svh_*(svh_ctx* ...)` declaration."""
    txt = open(os.path.join(ROOT, "include", "svh.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    out = []
    for m in re.finditer(r"\bint\s+(svh_\w+)\s*\(([^;{]*)\)\s*;", txt):
        params = [p.strip() for p in m.group(2).replace("\n", " ").split(",")]
        if params and params[0].startswith(("svh_ctx*", "svh_ctx *")):
            out.append((m.group(1), params))
    return out


def _null_ctx_calls():
    lines = []
    for name, params in _context_entry_points():
        args = ["nullptr"] + ["nullptr" if "*" in p else "0" for p in params[1:]]
        lines.append(f"CHECK({name}({', '.join(args)}) == SVH_ERR_ARG);")
    return "\n".join(lines) + "\n"


def write_null_ctx_calls(path):
    text = _null_ctx_calls()
    if not os.path.exists(path) or open(path).read() != text:
        with open(path, "w") as f:
            f.write(text)


def test_generated_calls_cover_context_entry_points():
    names = [n for n, _ in _context_entry_points()]
    assert "svh_vecchia_eval" in names and "svh_predict_noalloc" in names
    assert len(names) >= 30


@pytest.mark.skipif(not os.path.exists(os.path.join(SAN, "Makefile")),
                    reason="sanitizer sources not present")
@pytest.mark.skipif(shutil.which("gcc") is None or not os.path.exists("/opt/rocm/bin/hipcc"),
                    reason="gcc / hipcc not available")
def test_oracle_and_host_abi_under_asan_ubsan():
    os.makedirs(os.path.join(SAN, "_build"), exist_ok=True)
    write_null_ctx_calls(os.path.join(SAN, "_build", "abi_null_ctx.inc"))
    r = subprocess.run(["make", "-s", "-j8", "-C", SAN, "run"], capture_output=True, text=True,
                       timeout=900)
    log = r.stdout + r.stderr
    assert r.returncode == 0, log[-4000:]
    assert "oracle_checks: 0 failures" in log, log[-4000:]
    assert "abi_checks: 0 failures" in log, log[-4000:]
    assert "ERROR: AddressSanitizer" not in log and "runtime error:" not in log, log[-4000:]
