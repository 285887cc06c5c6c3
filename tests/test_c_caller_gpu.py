"""The library driven from a plain C program (tests/c/c_caller.c).

The program is linked to libblasdsp.so and the system HIP runtime and runs in
a process of its own with no Python or torch, as a C caller of the reference's
kernels would: every reference-named entry point family, the grouped launch,
the split-FP16 form and the generator, each against the CPU oracle.  Its
serving-loop check is switched off and printed as SKIPPED, which the second
test reports as a skip.
"""
import pathlib
import subprocess

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
EXE = ROOT / "tests" / "c" / "c_caller"


def test_c_caller_builds():
    if not EXE.exists():
        pytest.skip("tests/c/c_caller is not built (make -C tests/c)")
    out = subprocess.run(["ldd", str(EXE)], capture_output=True, text=True, check=True).stdout
    assert "libblasdsp.so" in out and "not found" not in out, out


_run = {}


@pytest.mark.gpu
def test_c_caller_on_gpu():
    assert EXE.exists(), "tests/c/c_caller is not built (make -C tests/c)"
    r = subprocess.run([str(EXE)], capture_output=True, text=True, timeout=120)
    _run["stdout"] = r.stdout
    print(r.stdout)
    print(r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("PASSED")


@pytest.mark.gpu
def test_c_caller_skipped_checks():
    """Checks the program reports as SKIPPED (its serving-loop check is
    switched off) show up here as skipped, not as passed."""
    if "stdout" not in _run:
        pytest.skip("the C program did not run")
    skipped = [ln for ln in _run["stdout"].splitlines() if ln.startswith("SKIPPED")]
    if skipped:
        pytest.skip("; ".join(skipped))
