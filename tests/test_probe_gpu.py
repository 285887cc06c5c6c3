"""The measurement probes behind `bench.py --workload mfma` and the HBM read
roof (blasdsp_probe_mfma, blasdsp_probe_lds, blasdsp_probe_read): every form
and mode runs on a small grid, the results are what the inputs fix, and
arguments outside what each accepts are refused before any launch."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N_FORMS = 8  # bf16 16x16x32 / 32x32x16, f16, f32 16x16x4 / 32x32x2, f64, fp8 16x16x128 / 32x32x64


def test_mfma_probe_every_form(blas, dev):
    """Operand bytes all 0x3c (positive and finite in BF16, FP16, FP32, FP64
    and E4M3): every form's per-thread sums come out finite and positive."""
    import torch

    L, s = blas.lib(), torch.cuda.current_stream().cuda_stream
    src = torch.full((4096,), 0x3C3C3C3C, dtype=torch.int32, device=dev)  # 1024 words of 16 bytes
    blocks = 8
    for form in range(N_FORMS):
        out = torch.full((256 * blocks,), float("nan"), dtype=torch.float32, device=dev)
        assert L.blasdsp_probe_mfma(form, src.data_ptr(), out.data_ptr(), blocks, 3, s) == 0, form
        torch.cuda.synchronize()
        o = out.cpu().numpy()
        assert np.isfinite(o).all() and (o > 0).all(), (form, o[:4])
    out = torch.empty(256, dtype=torch.float32, device=dev)
    for form in (-1, N_FORMS):
        assert L.blasdsp_probe_mfma(form, src.data_ptr(), out.data_ptr(), 1, 1, s) == blas.ERR_INVALID, form
    assert L.blasdsp_probe_mfma(0, src.data_ptr() + 4, out.data_ptr(), 1, 1, s) == blas.ERR_INVALID
    assert L.blasdsp_probe_mfma(0, src.data_ptr(), out.data_ptr(), 0, 1, s) == blas.ERR_INVALID
    assert L.blasdsp_probe_mfma(0, src.data_ptr(), out.data_ptr(), 1, 0, s) == blas.ERR_INVALID


def test_lds_probe_every_mode(blas, dev):
    """The load probe reads its 2 MiB region in every mode; with registers
    only it XORs words 0 and 3 of each vector, which are equal here."""
    import torch

    L, s = blas.lib(), torch.cuda.current_stream().cuda_stream
    region = torch.ones(1 << 19, dtype=torch.float32, device=dev)  # 2 MiB, the probe's extent
    blocks = 16
    for mode in range(3):
        out = torch.full((256 * blocks,), -1, dtype=torch.int32, device=dev)
        assert L.blasdsp_probe_lds(mode, region.data_ptr(), out.data_ptr(), blocks, 4, s) == 0, mode
        torch.cuda.synchronize()
        if mode == 2:
            assert (out.cpu().numpy() == 0).all()
    out = torch.empty(256, dtype=torch.int32, device=dev)
    assert L.blasdsp_probe_lds(3, region.data_ptr(), out.data_ptr(), 1, 1, s) == blas.ERR_INVALID
    assert L.blasdsp_probe_lds(0, region.data_ptr(), out.data_ptr(), 0, 1, s) == blas.ERR_INVALID


def test_read_probe_visits_every_word_once(blas, dev):
    """The per-workgroup XORs of the read probe combine to the XOR of every
    32-bit word of the buffer (each 16-byte vector is read exactly once)."""
    import torch

    L, s = blas.lib(), torch.cuda.current_stream().cuda_stream
    rng = np.random.default_rng(9)
    for nvec, blocks in ((1, 1), (1000, 3), (123457, 64)):
        words = rng.integers(0, 2**32, 4 * nvec, dtype=np.uint64).astype(np.uint32)
        buf = torch.from_numpy(words.view(np.int32)).to(dev)
        out = torch.zeros(blocks, dtype=torch.int32, device=dev)
        assert L.blasdsp_probe_read(buf.data_ptr(), 16 * nvec, out.data_ptr(), blocks, s) == 0
        torch.cuda.synchronize()
        got = np.bitwise_xor.reduce(out.cpu().numpy().view(np.uint32))
        assert got == np.bitwise_xor.reduce(words), (nvec, blocks)
    buf = torch.zeros(64, dtype=torch.int32, device=dev)
    out = torch.zeros(1, dtype=torch.int32, device=dev)
    for bad in ((buf.data_ptr(), 8), (buf.data_ptr(), 24), (buf.data_ptr() + 4, 64)):
        assert L.blasdsp_probe_read(bad[0], bad[1], out.data_ptr(), 1, s) == blas.ERR_INVALID, bad
