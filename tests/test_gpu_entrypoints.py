"""Every piecewise entry point of the C ABI against the CPU FP64 oracle, the
FP32 particle path (the benchmarked one) at small sizes, DoFixing /
DoInversion, the Cauchy invariant of the device fields, the device PLT
tables' cubic symmetry and the context destruction order.

Entry points (reference interface in parentheses):
  mfic_lpt_component   (ic_generator.cc:931-970, 987-1047)
  mfic_lattice_readout (particle_generator.hh:499-544, 597-639)
  mfic_grid_shift      (grid_fft.hh:832-842)
  mfic_grid_inverse_laplacian (grid_fft.hh:440-449)
  mfic_grid_zero_dc    (grid_fft.hh:844-850)
  mfic_grid_scale      (grid_fft.hh:400-412)
  mfic_lpt_sources     (ic_generator.cc:523-629)
Tolerances: FP64 rel-L2 1e-10, FP32 rel-L2 1e-5 (max error over the largest
value 1e-4).
"""
import ctypes as C

import numpy as np
import pytest

from monofonic_amd import F32, F64, Context, Grid, _lib as L, lpt
from oracle import oracle as O

from _gpu_helpers import lattice_q, max_rel, read_rows, rel_l2

TOL = {F64: 1e-10, F32: 1e-5}
TOL_MAX = {F64: 1e-9, F32: 1e-4}


def _up(ctx, arr, prec, box, space):
    n = arr.shape[0]
    g = Grid(ctx, (n,) * 3, (box,) * 3, prec=prec)
    g.upload(arr if prec == F64 else arr.astype(np.float32))
    g.set_space(space)
    return g


# ------------------------------------------------------ FP32 particle path --
@pytest.mark.gpu
@pytest.mark.parametrize("n", [32, 64])
@pytest.mark.parametrize("lattice,plt", [(L.SC, False), (L.SC, True), (L.BCC, False), (L.BCC, True)])
def test_fp32_particles_match_oracle(ctx, n, lattice, plt):
    """FP32 bulk, FP64 q+Psi: Psi = pos - q and v against the oracle's FP64
    particles (3LPT)"""
    box, z, seed = 1.5 * n, 24.0, 4242 + n
    print(f"FP32 particles n={n} lattice={lattice} plt={plt}", flush=True)
    cfg = O.ic_cfg(n, box, 3, seed, lattice=lattice, plt=int(plt))
    _, pos, vel, _ = O.ic_run(cfg, O.baseline_cosmology(z), particles=True)
    run = lpt.ICRun(ctx, lpt.ICConfig(n=n, box=box, zstart=z, lpt_order=3, lattice=lattice, plt=plt, prec=F32,
                                      seed=seed, particles=True))
    run.execute()
    p, v = run.particles()
    run.close()
    pos, vel = pos.reshape(-1, 3), vel.reshape(-1, 3)
    q = lattice_q(n, lattice, np.arange(len(pos)))
    e_psi, e_v = rel_l2(p - q, pos - q), rel_l2(v, vel)
    m_psi, m_v = max_rel(p - q, pos - q), max_rel(v, vel)
    print(f"  Psi rel-L2 {e_psi:.2e} max {m_psi:.2e}; v rel-L2 {e_v:.2e} max {m_v:.2e}")
    assert e_psi < TOL[F32] and e_v < TOL[F32]
    assert m_psi < TOL_MAX[F32] and m_v < TOL_MAX[F32]


# ------------------------------------------------- DoFixing / DoInversion --
@pytest.mark.gpu
@pytest.mark.parametrize("prec", [F64, F32])
@pytest.mark.parametrize("fixing,inversion", [(True, False), (False, True), (True, True)])
def test_fixing_inversion_match_oracle(ctx, prec, fixing, inversion):
    """noise amplitude fixing and phase inversion (ic_generator.cc:378-383)"""
    n, box, z, seed = 32, 100.0, 24.0, 12345
    cfg = O.ic_cfg(n, box, 3, seed, fixing=int(fixing), inversion=int(inversion))
    fields, _, _, _ = O.ic_run(cfg, O.baseline_cosmology(z))
    run = lpt.ICRun(ctx, lpt.ICConfig(n=n, box=box, zstart=z, lpt_order=3, prec=prec, seed=seed, fixing=fixing,
                                      inversion=inversion))
    run.execute()
    for c in range(6):
        err = rel_l2(run.field(c).download_r(), O.rview(fields[c]))
        assert err < TOL[prec], (c, err)
    run.close()


# ------------------------------------------------------- mfic_lpt_component --
@pytest.mark.gpu
@pytest.mark.parametrize("prec", [F64, F32])
@pytest.mark.parametrize("lattice,plt", [(L.SC, False), (L.BCC, True)])
def test_lpt_component_matches_oracle(ctx, prec, lattice, plt):
    """the per-component assembly kernel on the oracle's potentials: with
    growth factors 1 the device applies exactly the oracle's formula"""
    n, box, z, seed = 32, 80.0, 24.0, 99
    cfg = O.ic_cfg(n, box, 3, seed, lattice=lattice, plt=int(plt), lunit=2.0, vunit=3.0)
    oc = O.baseline_cosmology(z)
    pots = O.potentials(cfg, oc, O.ngenic_fill(seed, n))
    coef = O.lpt_coefs(cfg, oc)
    oplt = O.plt_tables(lattice, 64) if plt else None
    k = L.LptCoef()
    for name, _ in L.LptCoef._fields_:
        setattr(k, name, getattr(coef, name))
    k.g1 = k.g2 = k.g3 = k.g3c = 1.0
    phi, phi2, phi3, A3 = pots
    dg = [_up(ctx, a, prec, box, L.KSPACE) for a in (phi, phi2, phi3, *A3)]
    dplt = lpt.PltTables(ctx, lattice, 64) if plt else None
    out = Grid(ctx, (n,) * 3, (box,) * 3, prec=prec)
    for comp in range(6):
        want = O.component_k(cfg, coef, pots, comp, plt=oplt)
        unit = (cfg.vunit if comp >= 3 else cfg.lunit) / box
        lpt.lpt_component(ctx, 3, comp, k, (dg[0], dg[1], dg[2], dg[3:]), out, unit, plt=dplt)
        err = rel_l2(out.download_k(), O.kview(want))
        assert err < TOL[prec], (comp, err)
    if oplt is not None:
        O.lib().orc_plt_free(C.byref(oplt))
        dplt.close()
    for g in dg + [out]:
        g.close()


# ------------------------------------------------------ mfic_lpt_sources --
@pytest.mark.gpu
@pytest.mark.parametrize("prec", [F64, F32])
@pytest.mark.parametrize("order,n", [(2, 32), (3, 32), (3, 48)])
def test_lpt_sources_matches_oracle(ctx, prec, order, n):
    box, z, seed = 2.0 * n, 24.0, 31
    cfg = O.ic_cfg(n, box, order, seed)
    oc = O.baseline_cosmology(z)
    phi, phi2, phi3, A3 = O.potentials(cfg, oc, O.ngenic_fill(seed, n))
    coef = O.lpt_coefs(cfg, oc)
    g1 = _up(ctx, phi / coef.g1, prec, box, L.KSPACE)
    outs = [Grid(ctx, (n,) * 3, (box,) * 3, prec=prec, space=L.KSPACE) for _ in range(5)]
    if order == 2:
        lpt.lpt_sources(ctx, 2, g1, outs[0])
        checks = [(outs[0], phi2, coef.g2)]
    else:
        lpt.lpt_sources(ctx, 3, g1, outs[0], outs[1], outs[2:])
        checks = [(outs[0], phi2, coef.g2), (outs[1], phi3, coef.g3)] + \
                 [(outs[2 + d], A3[d], coef.g3c) for d in range(3)]
    for g, want, s in checks:
        err = rel_l2(g.download_k() * s, O.kview(want))
        assert err < (5e-9 if prec == F64 else TOL[prec]), err
    for g in outs + [g1]:
        g.close()


# --------------------------------------------------------- readout / shift --
@pytest.mark.gpu
@pytest.mark.parametrize("prec", [F64, F32])
@pytest.mark.parametrize("lattice", [L.SC, L.BCC])
@pytest.mark.parametrize("is_pos", [1, 0])
def test_lattice_readout_matches_oracle(ctx, prec, lattice, is_pos):
    torch = pytest.importorskip("torch")
    n, box, lunit = 32, 50.0, 3.0
    rng = np.random.default_rng(11 + lattice)
    f = O.empty_field(n)
    O.rview(f)[:] = rng.standard_normal((n, n, n)) * 0.01
    if prec == F32:
        f = f.astype(np.float32).astype(np.float64)
    for idim in range(3):
        g = _up(ctx, f, prec, box, L.RSPACE)
        npart = (1 << lattice) * n ** 3
        dst = torch.zeros(3 * npart, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        L.call("mfic_lattice_readout", g.h, lattice, idim, is_pos, lunit, C.c_void_p(dst.data_ptr()))
        ctx.sync()
        got = dst.cpu().numpy().reshape(-1, 3)[:, idim]
        want = O.lattice_readout(f.copy(), box, lattice, idim, bool(is_pos), lunit)[:, idim]
        if is_pos:
            q = lattice_q(n, lattice, np.arange(npart), lunit)[:, idim]
            got, want = got - q, want - q
        err = rel_l2(got, want)
        assert err < TOL[prec], (idim, err)
        g.close()


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [F64, F32])
@pytest.mark.parametrize("s", [(0.5, 0.5, 0.5), (0.25, -0.5, 1.0)])
def test_grid_shift_matches_oracle(ctx, prec, s):
    n, box = 24, 30.0
    rng = np.random.default_rng(5)
    f = O.empty_field(n)
    O.rview(f)[:] = rng.standard_normal((n, n, n))
    g = _up(ctx, f, prec, box, L.RSPACE)
    g.shift_field(s, True)
    want = f.copy()
    O.shift_field(want, box, s)
    assert g.space == L.RSPACE
    err = rel_l2(g.download_r(), O.rview(want))
    assert err < TOL[prec], err
    g.close()


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [F64, F32])
def test_kspace_helpers_match_oracle(ctx, prec):
    """apply_InverseLaplacian, zero_DC_mode and operator*= on a k-space grid"""
    n, box = 32, 40.0
    wn = O.ngenic_fill(3, n)
    O.kview(wn)[0, 0, 0] = 2.5 - 1.0j  # a DC value to clear
    for name in ("inverse_laplacian", "zero_dc", "scale"):
        g = _up(ctx, wn, prec, box, L.KSPACE)
        want = wn.copy()
        if name == "inverse_laplacian":
            g.apply_InverseLaplacian()
            O.inverse_laplacian(want, box)
        elif name == "zero_dc":
            g.zero_DC_mode()
            O.zero_dc(want)
        else:
            g *= -0.375
            O.scale(want, -0.375)
        got = g.download_k()
        assert rel_l2(got, O.kview(want)) < TOL[prec], name
        assert got[0, 0, 0] == O.kview(want)[0, 0, 0] or prec == F32
        g.close()


# --------------------------------------------------------- Cauchy invariant --
def _grad(n, L_, axis):
    m = np.arange(n if axis < 2 else n // 2 + 1, dtype=float)
    g = np.where(m > n // 2, m - n, m) * (2 * np.pi / L_)
    g[m == n // 2] = 0.0
    shape = [1, 1, 1]
    shape[axis] = -1
    return g.reshape(shape)


@pytest.mark.gpu
def test_cauchy_invariant_of_device_fields(ctx):
    """I_ij (testing.cc:221-268) formed from the device's 3LPT fields with the
    device convolver's gradient products matches the oracle's I, and is far
    below the third-order residual that a sign-flipped A3 curl leaves in the
    oracle's fields (tests/test_oracle_invariants.py pins the orders)"""
    n, box, z, seed = 32, 64.0, 99.0, 2024
    cfg = O.ic_cfg(n, box, 3, seed, lunit=box, vunit=1.0)
    oc = O.baseline_cosmology(z)
    pots = O.potentials(cfg, oc, O.ngenic_fill(seed, n))
    coef = O.lpt_coefs(cfg, oc)
    ref = [O.component_k(cfg, coef, pots, c) for c in range(6)]
    wrong = [O.component_k(cfg, coef, (pots[0], pots[1], pots[2], tuple(-a for a in pots[3])), c) for c in range(6)]

    def oracle_invariant(f, i, j):
        w = O.empty_field(n)
        O.kview(w)[:] = 1j * _grad(n, box, j) * O.kview(f[3 + i]) - 1j * _grad(n, box, i) * O.kview(f[3 + j])
        for k in range(3):
            O.convolve2(n, box, O.operand(f[k], O.GRAD, (i,)), O.operand(f[3 + k], O.GRAD, (j,)), w, O.ADD)
        for k in range(3):
            O.convolve2(n, box, O.operand(f[k], O.GRAD, (j,)), O.operand(f[3 + k], O.GRAD, (i,)), w, O.SUB)
        return O.kview(w)
    run = lpt.ICRun(ctx, lpt.ICConfig(n=n, box=box, zstart=z, lpt_order=3, prec=F64, seed=seed, lunit=box,
                                      vunit=1.0))
    run.execute()
    comps = []
    for c in range(6):
        g = Grid(ctx, (n,) * 3, (box,) * 3, prec=F64)
        g.copy_from(run.field(c))
        g.FourierTransformForward()
        comps.append(g)
    conv = lpt.OrszagConvolver(ctx, (n,) * 3, (box,) * 3, F64)
    inv = Grid(ctx, (n,) * 3, (box,) * 3, prec=F64, space=L.KSPACE)
    for i in range(3):
        j = (i + 1) % 3
        vi, vj = comps[3 + i].download_k(), comps[3 + j].download_k()
        inv.upload_k(1j * _grad(n, box, j) * vi - 1j * _grad(n, box, i) * vj)
        for k in range(3):
            conv.convolve_Gradients(comps[k], (i,), comps[3 + k], (j,), lpt.add_to(inv))
        for k in range(3):
            conv.convolve_Gradients(comps[k], (j,), comps[3 + k], (i,), lpt.subtract_from(inv))
        got = inv.download_k().copy()
        want = oracle_invariant(ref, i, j)
        bad = oracle_invariant(wrong, i, j)
        # I is a near-cancellation of terms far larger than itself: compare
        # the difference with the size of one velocity-gradient term
        scale_ = np.linalg.norm(vi * _grad(n, box, j))
        print(f"I_{i}{j}: |I_dev - I_oracle| / |dv_i/dq_j| = {np.linalg.norm(got - want) / scale_:.2e}, "
              f"|I_dev| / |I_wrong A3| = {np.linalg.norm(got) / np.linalg.norm(bad):.2e}")
        assert np.linalg.norm(got - want) < 1e-12 * scale_
        assert np.linalg.norm(got) < 0.1 * np.linalg.norm(bad)
    for g in comps + [inv]:
        g.close()
    conv.close()
    run.close()


# --------------------------------------------------------- PLT symmetry ----
@pytest.mark.gpu
@pytest.mark.parametrize("lattice", [L.SC, L.BCC])
def test_device_plt_tables_cubic_symmetry(ctx, lattice):
    from test_oracle_invariants import check_cubic_symmetry
    t = lpt.PltTables(ctx, lattice, 32)
    tab = t.download()
    t.close()
    err = check_cubic_symmetry(tab, 32, lattice)
    print(f"device PLT tables lattice {lattice}: max cubic-symmetry violation {err:.2e}")
    assert err < 3e-3


# ------------------------------------------------ context destruction order --
@pytest.mark.gpu
def test_context_refuses_destroy_with_live_handles():
    """mfic_ctx_destroy returns MFIC_ERR_STATE while a grid, convolver, run or
    PLT table of the context is alive, and succeeds once they are gone"""
    c = Context(0)
    lib = L.load()
    g = Grid(c, (16,) * 3, (10.0,) * 3, prec=F32)
    conv = lpt.OrszagConvolver(c, (16,) * 3, (10.0,) * 3, F32)
    assert lib.mfic_ctx_destroy(c.h) == -5
    assert b"live handle" in lib.mfic_last_error()
    g.close()
    assert lib.mfic_ctx_destroy(c.h) == -5
    conv.close()
    run = lpt.ICRun(c, lpt.ICConfig(n=16, box=20.0, zstart=24.0, lpt_order=2, prec=F32, seed=1))
    assert lib.mfic_ctx_destroy(c.h) == -5
    run.execute()
    run.close()
    assert lib.mfic_ctx_destroy(c.h) == 0
    c.h = C.c_void_p()
