"""The C-ABI library loads, exports every function include/blasdsp.h declares,
and rejects bad arguments on the host (no GPU needed: argument checks run
before any HIP call)."""
import ctypes
import json
import pathlib

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent


def test_library_exports_every_declared_function(blas):
    L = blas.lib()
    names = blas.header_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    for expected in ("dgemm_NN", "dgemm_TT", "sgemm_NT_qxx", "sgemm_TN_qxx", "sgemm_null",
                     "irr_shortA_dgemm_TN_dbuf", "irr_shortA_dgemm_null", "irr_shortA_sgemm_TT",
                     "irr_shortA_sgemm_null", "irr_thinB_dgemm_NN_dbuf", "irr_thinB_dgemm_null",
                     "irr_thinB_sgemm_TT", "irr_thinB_sgemm_null", "blasdsp_irr_thinB_sgemm_TT",
                     "blasdsp_sgemm_NT", "blasdsp_group_exec", "blasdsp_gemm_noalloc",
                     "dgemm_null", "blasdsp_dgemm", "blasdsp_sgemm_ex", "blasdsp_fill_f32"):
        assert expected in names


def test_binding_declares_every_function_with_arguments(blas):
    """Every header function taking arguments has ctypes argument types in
    the Python binding (an undeclared one would pass floats wrongly)."""
    import re

    L = blas.lib()
    text = blas.HEADER.read_text()
    undeclared = []
    for name in blas.header_functions():
        m = re.search(r"\b%s\s*\(([^)]*)\)" % re.escape(name), text)
        takes_args = m is not None and m.group(1).strip() not in ("", "void")
        if takes_args and getattr(L, name).argtypes is None:
            undeclared.append(name)
    assert not undeclared, undeclared


def test_version_and_status_strings(blas):
    assert "gfx950" in blas.version()
    L = blas.lib()
    assert L.blasdsp_status_string(0) == b"success"
    assert L.blasdsp_status_string(-22) == b"invalid argument"


FAKE = ctypes.c_void_p(0x1000)  # never dereferenced: every case fails the host checks


@pytest.mark.parametrize("args", [
    # (transA, transB, M, N, K, lda, ldb, ldc)
    ("X", "N", 4, 4, 4, 4, 4, 4),   # bad mode
    ("N", "Q", 4, 4, 4, 4, 4, 4),
    ("N", "N", -1, 4, 4, 4, 4, 4),  # negative dims
    ("N", "N", 4, -1, 4, 4, 4, 4),
    ("N", "N", 4, 4, -1, 4, 4, 4),
    ("N", "N", 4, 4, 8, 7, 4, 4),   # lda < K
    ("T", "N", 9, 4, 4, 8, 4, 4),   # lda < M (A stored K x M)
    ("N", "N", 4, 9, 4, 4, 8, 9),   # ldb < N
    ("N", "T", 4, 4, 9, 9, 8, 4),   # ldb < K (B stored N x K)
    ("N", "N", 4, 9, 4, 4, 9, 8),   # ldc < N
])
def test_invalid_arguments_rejected(blas, args):
    ta, tb, M, N, K, lda, ldb, ldc = args
    L = blas.lib()
    for fn in (L.blasdsp_dgemm, L.blasdsp_sgemm):
        st = fn(ta.encode(), tb.encode(), M, N, K, 1.0, FAKE, lda, FAKE, ldb, 0.0, FAKE, ldc, None)
        assert st == blas.ERR_INVALID


def test_null_pointers_rejected(blas):
    L = blas.lib()
    st = L.blasdsp_dgemm(b"N", b"N", 4, 4, 4, 1.0, None, 4, FAKE, 4, 0.0, FAKE, 4, None)
    assert st == blas.ERR_INVALID
    st = L.blasdsp_dgemm(b"N", b"N", 4, 4, 4, 1.0, FAKE, 4, FAKE, 4, 0.0, None, 4, None)
    assert st == blas.ERR_INVALID
    st = L.dgemm_NN(4, 4, 4, 4, 4, 4, FAKE, None, FAKE, 1.0, 1.0, None)
    assert st == blas.ERR_INVALID
    st = L.blasdsp_irr_thinB_sgemm_NN(4, 4, 4, 4, 4, 4, FAKE, FAKE, None, 1.0, 1.0, None)
    assert st == blas.ERR_INVALID
    # the reference's packed forms: checked before anything is written
    assert L.irr_thinB_sgemm_NN(4, 4, 4, FAKE, FAKE, None, 1.0, 1.0, None) == blas.ERR_INVALID
    assert L.sgemm_NN_qxx(4, 4, 4, None, FAKE, FAKE, 1.0, 3.0, None) == blas.ERR_INVALID
    assert L.irr_shortA_dgemm_TT_dbuf(4, 4, 4, FAKE, None, FAKE, None) == blas.ERR_INVALID
    assert L.irr_shortA_sgemm_NT(-1, 4, 4, FAKE, FAKE, FAKE, None) == blas.ERR_INVALID


def test_empty_problems_are_noops(blas):
    # M == 0 or N == 0: nothing to do, nothing launched, pointers unused
    L = blas.lib()
    assert L.blasdsp_dgemm(b"N", b"N", 0, 5, 5, 1.0, None, 5, None, 5, 0.0, None, 5, None) == 0
    assert L.blasdsp_sgemm(b"T", b"T", 5, 0, 5, 1.0, None, 5, None, 5, 0.0, None, 1, None) == 0
    assert L.blasdsp_sgemm_TT(5, 5, 1, 0, 0, 5, None, None, None, 1.0, 0.0, None) == 0
    assert L.sgemm_TT_qxx(0, 5, 5, None, None, None, 1.0, 3.0, None) == 0
    assert L.irr_thinB_dgemm_NT_dbuf(5, 0, 5, None, None, None, None) == 0


def test_bad_algo_and_fill_arguments(blas):
    L = blas.lib()
    st = L.blasdsp_dgemm_ex(b"N", b"N", 4, 4, 4, 1.0, FAKE, 4, FAKE, 4, 0.0, FAKE, 4, 7, 0, None)
    assert st == blas.ERR_INVALID or st == blas.ERR_DEVICE  # unknown algo
    assert L.blasdsp_fill_f64(FAKE, 4, 4, 3, 0, 0, 1, 0.0, None) == blas.ERR_INVALID
    assert L.blasdsp_fill_f32(FAKE, 4, 4, 4, 0, 9, 1, 0.0, None) == blas.ERR_INVALID


def test_serve_rejects_bad_config(blas):
    cfg = blas.ServeConfig(0, 8, 2, 1, 0.0)  # zero budget
    res = (blas.RequestResult * 1)()
    req = (blas.Request * 1)(blas.Request(0, b"N", b"N", 4, 4, 4, 1.0, 0.0, 1, 0.0))
    assert blas.lib().blasdsp_serve(req, 1, cfg, res, None, None) == blas.ERR_INVALID
    cfg = blas.ServeConfig(1 << 20, 8, 0, 1, 0.0)  # no launch slot
    assert blas.lib().blasdsp_serve(req, 1, cfg, res, None, None) == blas.ERR_INVALID
    cfg = blas.ServeConfig(1 << 20, 16385, 2, 1, 0.0)  # batch beyond the fill launch's grid
    assert blas.lib().blasdsp_serve(req, 1, cfg, res, None, None) == blas.ERR_INVALID


def test_tune_table_load_query_save(blas, tmp_path):
    L = blas.lib()
    L.blasdsp_tune_clear()
    assert blas.tune_query(blas.FLOAT64, "N", "N", 4096, 4096, 4096) is None
    p = tmp_path / "tune.txt"
    p.write_text("blasdsp-tune-v1 1 0 0 12 12 12 1 128 1 2400.0\n"
                 "blasdsp-tune-v1 0 1 0 5 16 13 2 0 4 330.5\n")
    assert L.blasdsp_tune_load(str(p).encode()) == 0
    c = blas.tune_query(blas.FLOAT64, "N", "N", 4000, 4200, 4096)   # same log2 buckets
    assert (c.algo, c.tile, c.splitk) == (blas.ALGO_TILE, 128, 1)
    c = blas.tune_query(blas.FLOAT32, "T", "N", 32, 65536, 8192)
    assert (c.algo, c.splitk) == (blas.ALGO_SKINNY, 4)
    assert blas.tune_query(blas.FLOAT32, "N", "N", 32, 65536, 8192) is None
    q = tmp_path / "out.txt"
    assert L.blasdsp_tune_save(str(q).encode()) == 0
    assert sorted(q.read_text().split("\n")) == sorted(p.read_text().replace("2400.0", "2400.000")
                                                       .replace("330.5", "330.500").split("\n"))
    bad = tmp_path / "bad.txt"
    bad.write_text("blasdsp-tune-v1 1 0 0 12 12 12 1 96 1 1.0\n")  # tile 96 is not a kernel
    assert L.blasdsp_tune_load(str(bad).encode()) == blas.ERR_INVALID
    assert L.blasdsp_tune_load(str(tmp_path / "missing.txt").encode()) == blas.ERR_INVALID
    L.blasdsp_tune_clear()


def test_one_hip_runtime_per_process():
    """Loading the library and then importing torch maps a single HIP runtime."""
    import subprocess
    import sys

    code = ("import sys; sys.path.insert(0, %r)\n"
            "import __graft_entry__\n"
            "__graft_entry__.load_package().lib()\n"
            "import torch\n"
            "maps = open('/proc/self/maps').read().splitlines()\n"
            "import json; print(json.dumps(sorted({l.split()[-1] for l in maps if 'libamdhip64' in l})))\n") % str(ROOT)
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, check=True, timeout=300)
    libs = json.loads(out.stdout.strip().splitlines()[-1])
    assert len(libs) == 1, libs


def test_transpose_rejects_bad_arguments(blas):
    """Checked on the host before any HIP call."""
    L = blas.lib()
    buf = ctypes.create_string_buffer(1 << 12)
    a = ctypes.addressof(buf)
    assert L.blasdsp_transpose_s(-1, 4, a, 4, a + 2048, 4, None) == blas.ERR_INVALID
    assert L.blasdsp_transpose_s(4, 4, a, 3, a + 2048, 4, None) == blas.ERR_INVALID      # lda < N
    assert L.blasdsp_transpose_s(4, 4, a, 4, a + 2048, 3, None) == blas.ERR_INVALID      # ldb < M
    assert L.blasdsp_transpose_d(4, 4, None, 4, a, 4, None) == blas.ERR_INVALID
    assert L.blasdsp_transpose_s(8, 8, a, 8, a + 16, 8, None) == blas.ERR_INVALID        # overlapping
    assert L.transpose_float_32x32_dbuf(0, 3, None, None, None) == blas.SUCCESS          # empty: no launch
    assert L.transpose_A2B(2, 4, 4, 4, a, a + 2048, None) == blas.ERR_INVALID            # lda < N


def test_stencil_rejects_bad_arguments(blas):
    """Checked on the host before any HIP call."""
    L = blas.lib()
    buf = ctypes.create_string_buffer(1 << 14)
    a = ctypes.addressof(buf)
    b = a + 8192
    assert L.blasdsp_stencil(-1, 4, a, 6, b, 6, 1, None) == blas.ERR_INVALID
    assert L.blasdsp_stencil(4, 4, a, 5, b, 6, 1, None) == blas.ERR_INVALID        # lda < N + 2
    assert L.blasdsp_stencil(4, 4, a, 6, b, 5, 1, None) == blas.ERR_INVALID        # haloed ldb < N + 2
    assert L.blasdsp_stencil(4, 4, a, 6, b, 3, 0, None) == blas.ERR_INVALID        # packed ldb < N
    assert L.blasdsp_stencil(4, 4, None, 6, b, 6, 1, None) == blas.ERR_INVALID
    assert L.blasdsp_stencil(4, 4, a, 6, a + 8 * 20, 6, 1, None) == blas.ERR_INVALID  # overlaps A
    assert L.blasdsp_stencil(0, 4, None, 6, None, 6, 1, None) == blas.SUCCESS       # empty: no launch
    assert L.blasdsp_stencil_iterate(4, 4, a, b, 6, -1, None) == blas.ERR_INVALID
    assert L.blasdsp_stencil_iterate(4, 4, a, b, 6, 0, None) == blas.SUCCESS        # no sweep
    assert L.stencil_vector(-2, 4, a, b, None) == blas.ERR_INVALID
    assert L.stencil_scalar(4, 4, a, a + 8, None) == blas.ERR_INVALID               # overlaps A
    assert L.stencil_null(-1, 4, a, b, None) == blas.ERR_INVALID


def test_group_exec_names_and_argument_counts(blas):
    """The by-name launcher takes the reference launchers' names and word
    counts (scalars, then 3 pointers) and rejects anything else on the host."""
    L = blas.lib()
    words = (ctypes.c_uint64 * 11)(*([4] * 8 + [0, 0, 0]))  # only name and count checks reach it
    assert L.blasdsp_group_exec(b"launch_dgemm_XX", 7, 3, words, None) == blas.ERR_UNSUPPORTED
    assert L.blasdsp_group_exec(b"launch_dgemm_NN", 8, 3, words, None) == blas.ERR_INVALID   # TN has 8
    assert L.blasdsp_group_exec(b"launch_dgemm_TN", 7, 3, words, None) == blas.ERR_INVALID
    assert L.blasdsp_group_exec(b"launch_sgemm_NN", 4, 2, words, None) == blas.ERR_INVALID
    assert L.blasdsp_group_exec(b"launch_irr_thinB_dgemm_NN", 4, 3, words, None) == blas.ERR_INVALID  # 5 words
    assert L.blasdsp_group_exec(None, 4, 3, words, None) == blas.ERR_INVALID
    # dims 4 x 4 x 4 with NULL operand pointers: the argument checks refuse them
    nulls = (ctypes.c_uint64 * 7)(4, 4, 4, 0, 0, 0, 0)
    assert L.blasdsp_group_exec(b"launch_irr_shortA_sgemm_NN", 4, 3, nulls, None) == blas.ERR_INVALID
    huge = (ctypes.c_uint64 * 7)(2 ** 63, 4, 4, 0, 1, 1, 1)
    assert L.blasdsp_group_exec(b"launch_sgemm_NN", 4, 3, huge, None) == blas.ERR_INVALID


def test_noalloc_rejects_oversized_explicit_split(blas):
    """An explicit split-K whose partials exceed the static scratch array is
    refused before any device work (64 MiB: 16 slices of 2048 x 2048 FP64
    need 512 MiB)."""
    L = blas.lib()
    n = 2048
    d = blas.GemmDesc(blas.FLOAT64, b"N", b"N", n, n, n, 1.0, 0.0, FAKE, n, FAKE, n, FAKE, n)
    assert L.blasdsp_gemm_noalloc(ctypes.byref(d), blas.ALGO_TILE, 16, None) == blas.ERR_NOMEM
    assert L.blasdsp_gemm_noalloc(None, blas.ALGO_TILE, 1, None) == blas.ERR_INVALID
    bad = blas.GemmDesc(7, b"N", b"N", n, n, n, 1.0, 0.0, FAKE, n, FAKE, n, FAKE, n)
    assert L.blasdsp_gemm_noalloc(ctypes.byref(bad), blas.ALGO_TILE, 1, None) == blas.ERR_INVALID
    short = blas.GemmDesc(blas.FLOAT32, b"N", b"N", n, n, n, 1.0, 0.0, FAKE, n - 1, FAKE, n, FAKE, n)
    assert L.blasdsp_gemm_noalloc(ctypes.byref(short), blas.ALGO_AUTO, 0, None) == blas.ERR_INVALID


def test_probe_arguments_rejected(blas):
    """The measurement probes refuse unknown forms and modes, null or
    misaligned buffers and empty grids before any launch."""
    L = blas.lib()
    bad_src = ctypes.c_void_p(0x1004)  # not 16-byte aligned
    for form in (-1, 8):
        assert L.blasdsp_probe_mfma(form, FAKE, FAKE, 1, 1, None) == blas.ERR_INVALID, form
    assert L.blasdsp_probe_mfma(0, None, FAKE, 1, 1, None) == blas.ERR_INVALID
    assert L.blasdsp_probe_mfma(0, bad_src, FAKE, 1, 1, None) == blas.ERR_INVALID
    assert L.blasdsp_probe_mfma(7, FAKE, None, 1, 1, None) == blas.ERR_INVALID
    assert L.blasdsp_probe_mfma(7, FAKE, FAKE, 0, 1, None) == blas.ERR_INVALID
    assert L.blasdsp_probe_mfma(7, FAKE, FAKE, 1, 0, None) == blas.ERR_INVALID
    for mode in (-1, 3):
        assert L.blasdsp_probe_lds(mode, FAKE, FAKE, 1, 1, None) == blas.ERR_INVALID, mode
    assert L.blasdsp_probe_lds(0, None, FAKE, 1, 1, None) == blas.ERR_INVALID
    assert L.blasdsp_probe_lds(0, FAKE, FAKE, 1, 0, None) == blas.ERR_INVALID
    for ptr, nbytes in ((FAKE, 8), (FAKE, 24), (bad_src, 64), (None, 64)):
        assert L.blasdsp_probe_read(ptr, nbytes, FAKE, 1, None) == blas.ERR_INVALID, (ptr, nbytes)
    assert L.blasdsp_probe_read(FAKE, 64, FAKE, 0, None) == blas.ERR_INVALID
