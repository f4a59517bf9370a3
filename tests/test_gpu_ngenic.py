"""Device NGenIC white noise vs the CPU oracle (random_ngenic.cc:84-154):
bit-exact in FP64 (RANLUX stream, rejection loop, correctly rounded
log/sin/cos on both sides), exact rounding of the FP64 values in FP32
(including the FP32 fast path over 512^3 modes)."""
import numpy as np
import pytest

from monofonic_amd import F32, F64, Grid, lpt
from oracle import oracle as O


@pytest.mark.gpu
@pytest.mark.parametrize("n,seed", [(8, 1), (16, 12345), (32, 777), (64, 12345)])
def test_noise_bit_exact_fp64(ctx, n, seed):
    print(f"ngenic n={n} seed={seed}", flush=True)
    g = Grid(ctx, (n,) * 3, (100.0,) * 3, prec=F64)
    lpt.ngenic_fill(g, seed)
    got = g.download(np.float64)
    ref = O.ngenic_fill(seed, n)
    assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))
    g.close()


@pytest.mark.gpu
def test_noise_fp32_is_rounded_fp64(ctx):
    n, seed = 32, 4242
    g = Grid(ctx, (n,) * 3, (100.0,) * 3, prec=F32)
    lpt.ngenic_fill(g, seed)
    got = g.download(np.float32)
    ref = O.ngenic_fill(seed, n).astype(np.float32)
    assert np.array_equal(got, ref)
    g.close()


@pytest.mark.gpu
@pytest.mark.parametrize("n,seed", [(256, 12345), (512, 12345)])
def test_noise_fp32_fast_path_large(ctx, n, seed):
    """the FP32 fill takes a fast path for log/sin/cos with a rounding
    check; over 10^7-10^8 modes it must still equal the rounding of the
    correctly rounded FP64 fill (itself bit-exact with the oracle)"""
    print(f"ngenic FP32 vs FP64 n={n}", flush=True)
    g32 = Grid(ctx, (n,) * 3, (100.0,) * 3, prec=F32)
    lpt.ngenic_fill(g32, seed)
    got = g32.download(np.float32)
    g32.close()
    g64 = Grid(ctx, (n,) * 3, (100.0,) * 3, prec=F64)
    lpt.ngenic_fill(g64, seed)
    ref = g64.download(np.float64).astype(np.float32)
    g64.close()
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


@pytest.mark.gpu
def test_noise_structure(ctx):
    """zero Nyquist planes and DC, Hermitian kz=0 plane, unit mode power"""
    n = 64
    g = Grid(ctx, (n,) * 3, (100.0,) * 3, prec=F64)
    lpt.ngenic_fill(g, 99)
    k = g.download_k()
    h = n // 2
    assert np.all(k[h] == 0) and np.all(k[:, h] == 0) and np.all(k[:, :, h] == 0) and k[0, 0, 0] == 0
    i = np.arange(n)
    part = (-i) % n
    plane = k[:, :, 0]
    assert np.allclose(plane, np.conj(plane[part][:, part]))
    p = np.abs(k[np.abs(k) > 0]) ** 2
    assert abs(p.mean() - 1.0) < 0.01
    g.close()
