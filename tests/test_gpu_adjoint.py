"""Differentiable 2LPT on the device: the hand-written adjoint against
central finite differences (exact up to rounding for the quadratic 2LPT
map), a dot-product test, FP32 against FP64, and the PyTorch autograd
wrapper on a cloud-in-cell density loss."""
import numpy as np
import pytest

from monofonic_amd import F32, F64, Grid, lpt
from monofonic_amd.batch import MT19937_64

torch = pytest.importorskip("torch")
from monofonic_amd.adjoint import LPT2, cic_density, lpt2_map  # noqa: E402


def real_noise(ctx, n, box, seed):
    """NGenIC white noise (random_ngenic.cc:84-154) in real space, FP64"""
    g = Grid(ctx, (n,) * 3, (box,) * 3, prec=F64)
    lpt.ngenic_fill(g, seed)
    g.FourierTransformBackward()
    x = g.download_r().copy()
    g.close()
    return x


def cells(n, count, seed=7):
    g = MT19937_64(seed)
    return [g() % n ** 3 for _ in range(count)]


@pytest.mark.gpu
def test_adjoint_matches_central_differences(ctx):
    n, box = 32, 64.0
    plan = LPT2(ctx, n, box, prec=F64)
    d = torch.tensor(real_noise(ctx, n, box, 12345), device="cuda")
    C = torch.tensor(np.random.default_rng(3).standard_normal((3, n, n, n)), device="cuda")
    grad = plan.adjoint(d, C).reshape(-1)
    rms = grad.square().mean().sqrt().item()
    eps = 1e-3
    worst = 0.0
    for c in cells(n, 16):
        e = torch.zeros(n ** 3, dtype=torch.float64, device="cuda")
        e[c] = eps
        e = e.reshape(n, n, n)
        fd = ((plan.forward(d + e) * C).sum() - (plan.forward(d - e) * C).sum()).item() / (2 * eps)
        worst = max(worst, abs(fd - grad[c].item()) / rms)
    print(f"adjoint vs central differences (16 cells, FP64): max |diff| / rms(grad) = {worst:.2e}")
    assert worst < 1e-6
    # dot-product test along a random direction
    v = torch.tensor(np.random.default_rng(4).standard_normal((n, n, n)), device="cuda")
    lhs = ((plan.forward(d + eps * v) - plan.forward(d - eps * v)) * C).sum().item() / (2 * eps)
    rhs = (grad.reshape(n, n, n) * v).sum().item()
    print(f"dot-product test: {lhs:.12e} vs {rhs:.12e}")
    assert abs(lhs - rhs) <= 1e-9 * abs(rhs)
    plan.close()


@pytest.mark.gpu
def test_adjoint_fp32_vs_fp64(ctx):
    n, box = 64, 128.0
    x = real_noise(ctx, n, box, 12345)
    C = np.random.default_rng(5).standard_normal((3, n, n, n))
    out = {}
    for prec, dt in ((F64, torch.float64), (F32, torch.float32)):
        plan = LPT2(ctx, n, box, prec=prec)
        out[prec] = plan.adjoint(torch.tensor(x, dtype=dt, device="cuda"),
                                 torch.tensor(C, dtype=dt, device="cuda")).double().cpu().numpy()
        plan.close()
    err = np.linalg.norm(out[F32] - out[F64]) / np.linalg.norm(out[F64])
    print(f"FP32 adjoint vs FP64 adjoint: rel-L2 {err:.2e}")
    assert err < 1e-4


@pytest.mark.gpu
def test_autograd_cic_loss(ctx):
    n, box = 32, 64.0
    plan = LPT2(ctx, n, box, prec=F64)
    d0 = torch.tensor(real_noise(ctx, n, box, 12345), device="cuda")
    with torch.no_grad():
        target = cic_density(plan.forward(torch.tensor(real_noise(ctx, n, box, 54321), device="cuda")), box)

    def loss_of(d):
        return ((cic_density(lpt2_map(d, plan), box) - target) ** 2).sum()

    d = d0.clone().requires_grad_(True)
    loss = loss_of(d)
    loss.backward()
    grad = d.grad.reshape(-1)
    rms = grad.square().mean().sqrt().item()
    eps = 1e-3
    errs = []
    with torch.no_grad():
        for c in cells(n, 16):
            e = torch.zeros(n ** 3, dtype=torch.float64, device="cuda")
            e[c] = eps
            e = e.reshape(n, n, n)
            fd = (loss_of(d0 + e) - loss_of(d0 - e)).item() / (2 * eps)
            errs.append(abs(fd - grad[c].item()) / rms)
    print(f"CIC loss gradient vs central differences: median {np.median(errs):.2e}, max {max(errs):.2e} (of rms)")
    assert np.median(errs) < 1e-4 and max(errs) < 1e-2
    plan.close()


@pytest.mark.gpu
def test_c5_gradient_256_matches_central_differences(ctx):
    """the differentiable 2LPT workload at full size: 256^3, box 500 Mpc/h,
    z_start 24, loss = sum_cells (CIC density - target)^2 with the target from
    seed 54321 through the same map, noise from seed 12345; the FP64 adjoint
    gradient against FP64 central differences (eps 1e-3) on 64 cells drawn
    with mt19937_64(7), and the FP32 gradient against the FP64 one"""
    n, box = 256, 500.0
    print("C5 256^3 adjoint vs central differences", flush=True)
    plans = {F64: LPT2(ctx, n, box, prec=F64), F32: LPT2(ctx, n, box, prec=F32)}
    d64 = torch.tensor(real_noise(ctx, n, box, 12345), device="cuda")
    with torch.no_grad():
        tgt = cic_density(plans[F64].forward(torch.tensor(real_noise(ctx, n, box, 54321), device="cuda")), box)

    def loss_of(d, prec):
        t = tgt if prec == F64 else tgt.float()
        return ((cic_density(lpt2_map(d, plans[prec]), box) - t) ** 2).sum()

    grads = {}
    for prec in (F64, F32):
        d = (d64 if prec == F64 else d64.float()).clone().requires_grad_(True)
        loss_of(d, prec).backward()
        grads[prec] = d.grad.double().reshape(-1)
    g64 = grads[F64]
    rel32 = (torch.linalg.norm(grads[F32] - g64) / torch.linalg.norm(g64)).item()
    rms = g64.square().mean().sqrt().item()
    eps = 1e-3
    errs = []
    with torch.no_grad():
        for c in cells(n, 64):
            e = torch.zeros(n ** 3, dtype=torch.float64, device="cuda")
            e[c] = eps
            e = e.reshape(n, n, n)
            fd = (loss_of(d64 + e, F64) - loss_of(d64 - e, F64)).item() / (2 * eps)
            errs.append(abs(fd - g64[c].item()) / rms)
    print(f"C5: FD median {np.median(errs):.2e} max {max(errs):.2e} of rms(grad); FP32 vs FP64 gradient "
          f"rel-L2 {rel32:.2e}", flush=True)
    assert np.median(errs) < 1e-4 and max(errs) < 1e-2
    assert rel32 < 1e-3
    for p in plans.values():
        p.close()
