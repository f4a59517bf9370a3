"""The device complex FFT (mfic_fft_c2c) against numpy, and the device
semiclassical PT fields (mfic_ppt_fields, ic_generator.cc:786-897) against
the oracle on the same run.  FP64: rel-L2 1e-10 (FFT) / 1e-9 (fields);
FP32: 1e-5 (FFT) / 1e-4 (fields: the wave-function phases reach tens of
radians, so FP32 rounding of phi / (hbar D+) is amplified that much)."""
import ctypes as C

import numpy as np
import pytest

from monofonic_amd import F32, F64, Grid, _lib as L, lpt
from oracle import oracle as O

from _gpu_helpers import rel_l2

torch = pytest.importorskip("torch")


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [F64, F32])
@pytest.mark.parametrize("shape", [(16, 24, 32), (64, 64, 64), (9, 12, 48)])
def test_fft_c2c_matches_numpy(ctx, prec, shape):
    rng = np.random.default_rng(sum(shape))
    z = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    dt = torch.complex128 if prec == F64 else torch.complex64
    for sign, want in ((1, np.fft.fftn(z, norm="ortho")), (0, np.fft.ifftn(z, norm="ortho"))):
        t = torch.tensor(z, dtype=dt, device="cuda")
        torch.cuda.synchronize()
        L.call("mfic_fft_c2c", ctx.h, C.c_void_p(t.data_ptr()), (C.c_size_t * 3)(*shape), prec, sign)
        ctx.sync()
        err = rel_l2(t.cpu().numpy().astype(np.complex128), want)
        assert err < (1e-10 if prec == F64 else 1e-5), (sign, err)


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [F64, F32])
@pytest.mark.parametrize("order", [1, 2, 3])
def test_ppt_fields_match_oracle(ctx, prec, order):
    n, box, z, seed, vu = 32, 100.0, 49.0, 12345, 2.5
    print(f"PPT n={n} order={order} prec={prec}", flush=True)
    cfg = O.ic_cfg(n, box, order, seed)
    oc = O.baseline_cosmology(z)
    phi, phi2, _, _ = O.potentials(cfg, oc, O.ngenic_fill(seed, n))
    coef = O.lpt_coefs(cfg, oc)
    (want, hb_want) = O.ppt_fields(cfg, coef, phi, phi2, vu)
    run = lpt.ICRun(ctx, lpt.ICConfig(n=n, box=box, zstart=z, lpt_order=order, prec=prec, seed=seed))
    run.execute()
    outs = [Grid(ctx, (n,) * 3, (box,) * 3, prec=prec) for _ in range(4)]
    hb = C.c_double()
    L.call("mfic_ppt_fields", run.h, vu, *[g.h for g in outs], C.byref(hb))
    assert abs(hb.value / hb_want - 1) < (1e-12 if prec == F64 else 1e-5)
    for name, g, w in zip(("delta", "vx", "vy", "vz"), outs, want):
        err = rel_l2(g.download_r(), O.rview(w))
        print(f"  {name}: rel-L2 {err:.2e}")
        assert err < (1e-9 if prec == F64 else 1e-4), (name, err)
    for g in outs:
        g.close()
    run.close()


@pytest.mark.gpu
def test_ppt_full_size_is_unitary(ctx):
    """256^3 FP32: mean delta = 0 (the drift is unitary) and finite fields"""
    n, box = 256, 500.0
    run = lpt.ICRun(ctx, lpt.ICConfig(n=n, box=box, zstart=24.0, lpt_order=2, prec=F32, seed=12345))
    run.execute()
    outs = [Grid(ctx, (n,) * 3, (box,) * 3, prec=F32) for _ in range(4)]
    L.call("mfic_ppt_fields", run.h, 1.0, *[g.h for g in outs], None)
    d = outs[0].download_r()
    print(f"256^3 PPT: mean delta {d.mean():.2e}, std {d.std():.3f}")
    assert abs(d.mean()) < 1e-5 * d.std() * 100 and np.isfinite(d).all()
    for g in outs:
        assert np.isfinite(g.download_r()).all()
        g.close()
    run.close()
