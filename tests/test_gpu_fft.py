"""GPU parity of the unitary 3D r2c / c2r transforms (grid_fft.cc:207-270)
against numpy.fft.rfftn/irfftn(norm="ortho") (pocketfft, FP64)."""
import numpy as np
import pytest

from monofonic_amd import F32, F64, KSPACE, RSPACE, Grid

SHAPES = [(8, 8, 8), (12, 12, 12), (16, 16, 16), (24, 24, 24), (32, 32, 32), (48, 48, 48), (64, 64, 64),
          (96, 96, 96), (16, 32, 48), (72, 72, 72), (128, 128, 128), (144, 144, 144), (192, 192, 192),
          (32, 24, 256), (64, 64, 384), (288, 288, 288), (24, 16, 512), (16, 16, 768)]


def _roundtrip(ctx, shape, prec, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(shape)
    a = np.zeros(shape[:2] + (shape[2] + 2,))
    a[:, :, : shape[2]] = x
    g = Grid(ctx, shape, (100.0,) * 3, prec=prec)
    g.upload(a)
    g.FourierTransformForward()
    assert g.space == KSPACE
    k = g.download_k()
    ref = np.fft.rfftn(x, norm="ortho")
    g.FourierTransformBackward()
    assert g.space == RSPACE
    back = g.download_r()
    g.close()
    return k, ref, back, x


@pytest.mark.gpu
@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("prec,tol", [(F64, 1e-12), (F32, 2e-6)])
def test_fft_matches_numpy(ctx, shape, prec, tol):
    print(f"fft shape={shape} prec={prec}", flush=True)
    k, ref, back, x = _roundtrip(ctx, shape, prec, seed=sum(shape))
    rel = np.linalg.norm(k - ref) / np.linalg.norm(ref)
    assert rel < tol, rel
    rel_rt = np.linalg.norm(back - x) / np.linalg.norm(x)
    assert rel_rt < tol, rel_rt


@pytest.mark.gpu
@pytest.mark.parametrize("shape", [(768, 16, 16), (16, 768, 16), (512, 24, 16), (24, 512, 16)])
@pytest.mark.parametrize("prec,tol", [(F64, 1e-12), (F32, 2e-6)])
def test_fft_alternative_radix_plan(ctx, shape, prec, tol):
    """the strided passes on the alternative radix plan of their length"""
    from monofonic_amd import _lib as L
    from monofonic_amd import batch
    length = max(shape[:2])
    assert batch.plan_count(length) == 2
    print(f"fft shape={shape} prec={prec} plan 1 {batch.plan_info(length, 1)}", flush=True)
    batch.set_plan(ctx, length, L.PASS_PENCILS, prec, 1)
    try:
        assert batch.get_plan(ctx, length, L.PASS_PENCILS, prec) == 1
        k, ref, back, x = _roundtrip(ctx, shape, prec, seed=7)
    finally:
        batch.set_plan(ctx, length, L.PASS_PENCILS, prec, -1)
    assert batch.get_plan(ctx, length, L.PASS_PENCILS, prec) == 0
    assert np.linalg.norm(k - ref) / np.linalg.norm(ref) < tol
    assert np.linalg.norm(back - x) / np.linalg.norm(x) < tol


@pytest.mark.gpu
def test_plan_and_tile_keys(ctx):
    """the fused-row form and tile overrides read back as set, restore to the
    built-in choice, and out-of-range choices are refused"""
    import ctypes as C
    from monofonic_amd import _lib as L
    from monofonic_amd import batch
    batch.set_plan(ctx, 768, L.PASS_FUSED_ROWS, F32, 1)
    assert batch.get_plan(ctx, 768, L.PASS_FUSED_ROWS, F32) == 1
    assert batch.get_plan(ctx, 768, L.PASS_FUSED_ROWS, F64) == 0
    with pytest.raises(L.MficError):
        batch.set_plan(ctx, 768, L.PASS_FUSED_ROWS, F32, 2)
    batch.set_plan(ctx, 768, L.PASS_FUSED_ROWS, F32, -1)
    assert batch.get_plan(ctx, 768, L.PASS_FUSED_ROWS, F32) == 0
    t, pad = C.c_int(), C.c_int()
    L.call("mfic_fft_set_tile", ctx.h, 768, L.PASS_TRUNC_AX1, F32, 4, 0)
    L.call("mfic_fft_get_tile", ctx.h, 768, L.PASS_TRUNC_AX1, F32, C.byref(t), C.byref(pad))
    assert (t.value, pad.value) == (4, 0)
    assert batch.chosen_tiles(ctx, 512, F32)[(768, L.PASS_TRUNC_AX1)] == 4
    L.call("mfic_fft_set_tile", ctx.h, 768, L.PASS_TRUNC_AX1, F32, 0, 0)
    L.call("mfic_fft_get_tile", ctx.h, 768, L.PASS_TRUNC_AX1, F32, C.byref(t), C.byref(pad))
    assert t.value == 0 and (768, L.PASS_TRUNC_AX1) not in batch.chosen_tiles(ctx, 512, F32)
    with pytest.raises(L.MficError):
        L.call("mfic_fft_set_tile", ctx.h, 768, L.PASS_TRUNC_AX1, F32, 3, 0)


@pytest.mark.gpu
def test_fft_single_mode(ctx):
    """a single cosine mode maps to two half-amplitude delta peaks"""
    n = 64
    i = np.arange(n)
    x = np.cos(2 * np.pi * 3 * i / n)[:, None, None] * np.ones((n, n, n))
    a = np.zeros((n, n, n + 2))
    a[:, :, :n] = x
    g = Grid(ctx, (n, n, n), (1.0,) * 3, prec=F64)
    g.upload(a)
    g.FourierTransformForward()
    k = g.download_k()
    peak = 0.5 * np.sqrt(n ** 3)
    assert abs(k[3, 0, 0] - peak) < 1e-9 and abs(k[n - 3, 0, 0] - peak) < 1e-9
    k[3, 0, 0] = k[n - 3, 0, 0] = 0
    assert np.abs(k).max() < 1e-9
