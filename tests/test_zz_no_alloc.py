"""The no-allocation phi1 entry point (mfic_phi1_from_noise_noalloc) against
mfic_phi1_from_noise through the C ABI: bit-identical results on two calls in
a row (the second on a smaller grid, so stale table entries of the first
would show), no device memory taken by the call, and the documented refusals
(too large a grid, a tabulated transfer function)."""
import ctypes as C

import numpy as np
import pytest

from monofonic_amd import F32, F64, Context, Grid, _lib as L, lpt


def _mem_free():
    hip = lpt._hip()
    hip.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
    f, t = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(f), C.byref(t)) == 0
    return f.value


def _phi(ctx, n, box, prec, seed, noalloc, cos):
    g = Grid(ctx, (n,) * 3, (box,) * 3, prec=prec)
    lpt.ngenic_fill(g, seed)
    name = "mfic_phi1_from_noise_noalloc" if noalloc else "mfic_phi1_from_noise"
    L.call(name, g.h, g.h, C.byref(cos), 0, 0)
    ctx.sync()
    out = g.download_k().copy()
    g.close()
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [F64, F32])
def test_noalloc_phi1_matches_allocating_entry_point(prec):
    ctx = Context(0)
    try:
        cos = lpt.baseline_cosmology(24.0)
        for n, box, seed in ((96, 150.0, 12345), (48, 60.0, 777)):
            print(f"no-alloc phi1 n={n} prec={prec}", flush=True)
            want = _phi(ctx, n, box, prec, seed, False, cos)
            got = _phi(ctx, n, box, prec, seed, True, cos)
            assert np.array_equal(got, want), n
    finally:
        ctx.close()


@pytest.mark.gpu
def test_noalloc_phi1_takes_no_device_memory():
    ctx = Context(0)
    try:
        cos = lpt.baseline_cosmology(24.0)
        n = 256
        g = Grid(ctx, (n,) * 3, (500.0,) * 3, prec=F32)
        lpt.ngenic_fill(g, 1)
        L.call("mfic_phi1_from_noise_noalloc", g.h, g.h, C.byref(cos), 0, 0)  # first launch loads the kernels
        ctx.sync()
        lpt.ngenic_fill(g, 2)
        ctx.sync()
        before = _mem_free()
        L.call("mfic_phi1_from_noise_noalloc", g.h, g.h, C.byref(cos), 0, 0)
        ctx.sync()
        after = _mem_free()
        print(f"free device memory before / after: {before} / {after}", flush=True)
        assert after >= before
        g.close()
    finally:
        ctx.close()


@pytest.mark.gpu
def test_noalloc_phi1_refuses_what_it_cannot_do():
    ctx = Context(0)
    try:
        lib = L.load()
        cos = lpt.baseline_cosmology(24.0)
        g = Grid(ctx, (1024,) * 3, (1000.0,) * 3, prec=F32, space=L.KSPACE)
        assert lib.mfic_phi1_from_noise_noalloc(g.h, g.h, C.byref(cos), 0, 0) == -2
        assert b"no-allocation scratch" in lib.mfic_last_error()
        g.close()
        k = np.logspace(-5, 3, 400)
        T = 1.0 / (1.0 + (k / 0.05) ** 2)
        tab = lpt.baseline_cosmology(24.0)
        tab.tk_k = k.ctypes.data_as(C.POINTER(C.c_double))
        tab.tk_T = T.ctypes.data_as(C.POINTER(C.c_double))
        tab.tk_n = len(k)
        g = Grid(ctx, (32,) * 3, (100.0,) * 3, prec=F32, space=L.KSPACE)
        assert lib.mfic_phi1_from_noise_noalloc(g.h, g.h, C.byref(tab), 0, 0) == -7
        g.close()
    finally:
        ctx.close()
