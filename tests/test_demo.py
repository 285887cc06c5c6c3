"""The C demo (blas-for-multi-core-dsp_amd/demo/gemm_demo.c): the reference's
demo command line (family dev_id M K N num_threads) on the library."""
import pathlib
import subprocess

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
EXE = ROOT / "blas-for-multi-core-dsp_amd" / "demo" / "gemm_demo"


def test_demo_usage_without_gpu():
    """Too few arguments: prints the usage and returns before any HIP call."""
    if not EXE.exists():
        pytest.skip("demo not built (make -C blas-for-multi-core-dsp_amd/demo)")
    r = subprocess.run([str(EXE)], capture_output=True, text=True, timeout=60)
    assert "Usage" in r.stdout and r.returncode != 0


@pytest.mark.gpu
@pytest.mark.parametrize("args", [["dgemm", "0", "1512", "2048", "1536", "8"],
                                  ["sgemm", "0", "300", "257", "333", "1"],
                                  ["irr_shortA_dgemm", "0", "24", "4096", "3000", "1"],
                                  ["irr_thinB_sgemm", "0", "5000", "700", "32", "1"]])
def test_demo_runs_and_modes_agree(args):
    assert EXE.exists(), "demo not built"
    r = subprocess.run([str(EXE)] + args, capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Success!" in r.stdout
    for mode in ("NN", "NT", "TN", "TT"):
        line = next(ln for ln in r.stdout.splitlines() if ln.startswith(f"{args[0]}_{mode} elapsed time"))
        assert float(line.split("GFLOPS:")[1]) > 0, line
    if args[0] in ("dgemm", "sgemm"):
        assert "bit-identical" in r.stdout
