"""fcc particle runs and baryon + CDM particle runs on the device (ic_generator.cc:704-778 with the
Eisenstein & Hu fit, where delta_bc = theta_bc = 0): the dark matter is the
single-species run, the baryons read the same displacement and velocity
fields out on the secondary lattice (particle_generator.hh:58-63, 499-544,
597-639), both against the oracle's readout of lattice index 0 and 1; the
indexed readout entry point against the oracle; and both particle files
(Gadget-HDF5 with per-particle masses, SWIFT with the gas fields) with the
baryons' IDs after the dark matter's.

Tolerances: FP64 rel-L2 1e-10 (max over the largest value 1e-9), FP32 rel-L2
1e-5 (1e-4)."""
import ctypes as C

import numpy as np
import pytest

from monofonic_amd import F32, F64, Grid, _lib as L, lpt
from oracle import oracle as O

from _gpu_helpers import SUBLATTICES, max_rel, rel_l2

TOL = {F64: 1e-10, F32: 1e-5}
TOL_MAX = {F64: 1e-9, F32: 1e-4}
SECOND = {L.SC: (0.5, 0.5, 0.5), L.BCC: (0.5, 0.5, 0.0), L.FCC: (0.5, 0.5, 0.5)}


def species_q(n, lattice, lunit, index):
    """Lagrangian positions of lattice index 0 or 1 (the secondary lattice),
    sub-lattice s at (ijk + shift_s + second) / n * lunit, not wrapped"""
    i, j, k = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    ijk = np.stack([i, j, k]).reshape(3, -1).T.astype(np.float64)
    g = np.array(SECOND[lattice]) if index else np.zeros(3)
    return np.concatenate([(ijk + np.array(sub) + g) / n * lunit for sub in SUBLATTICES[lattice]])


def oracle_species(cfg, cos):
    """dark matter and baryon particles of the oracle run: the component
    fields read out on lattice index 0 and 1"""
    fields, pos, vel, _ = O.ic_run(cfg, cos, particles=True)
    n = cfg.n
    bp = np.zeros(((1 << cfg.lattice) * n ** 3, 3))
    bv = np.zeros_like(bp)
    for comp in range(6):
        f = fields[comp].copy()
        if comp < 3:
            bp[:, comp] = O.lattice_readout_index(f, cfg.L, cfg.lattice, 1, comp, True, cfg.lunit)[:, comp]
        else:
            bv[:, comp - 3] = O.lattice_readout_index(f, cfg.L, cfg.lattice, 1, comp - 3, False)[:, comp - 3]
    return pos.reshape(-1, 3), vel.reshape(-1, 3), bp, bv


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [F64, F32])
@pytest.mark.parametrize("plt", [False, True])
def test_fcc_run_matches_oracle(ctx, prec, plt):
    """ParticleLoad fcc: four sub-lattices (0,0,0), (0,1/2,1/2), (1/2,0,1/2),
    (1/2,1/2,0), with and without the fcc PLT corrections, 3LPT"""
    n, box, z, seed = 32, 48.0, 24.0, 1618
    print(f"fcc: n={n} plt={plt} prec={prec}", flush=True)
    cfg = O.ic_cfg(n, box, 3, seed, lattice=L.FCC, plt=int(plt), lunit=box, vunit=2.0)
    _, pos, vel, _ = O.ic_run(cfg, O.baseline_cosmology(z), particles=True)
    run = lpt.ICRun(ctx, lpt.ICConfig(n=n, box=box, zstart=z, lpt_order=3, lattice=L.FCC, plt=plt, prec=prec,
                                      seed=seed, particles=True, lunit=box, vunit=2.0))
    run.execute()
    p, v = run.particles()
    run.close()
    assert len(p) == 4 * n ** 3
    q = species_q(n, L.FCC, box, 0)
    pos, vel = pos.reshape(-1, 3), vel.reshape(-1, 3)
    e, m = rel_l2(p - q, pos - q), max_rel(p - q, pos - q)
    ev, mv = rel_l2(v, vel), max_rel(v, vel)
    print(f"  Psi rel-L2 {e:.2e} max {m:.2e}; v rel-L2 {ev:.2e} max {mv:.2e}", flush=True)
    assert e < TOL[prec] and m < TOL_MAX[prec] and ev < TOL[prec] and mv < TOL_MAX[prec]


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [F64, F32])
@pytest.mark.parametrize("lattice,plt,order", [(L.SC, False, 2), (L.BCC, False, 3), (L.BCC, True, 3),
                                               (L.FCC, False, 2)])
def test_baryon_run_matches_oracle(ctx, prec, lattice, plt, order):
    n, box, z, seed = 32, 48.0, 24.0, 2718
    print(f"baryons: n={n} lattice={lattice} plt={plt} order={order} prec={prec}", flush=True)
    cfg = O.ic_cfg(n, box, order, seed, lattice=lattice, plt=int(plt), lunit=box, vunit=2.0)
    dpos, dvel, bpos, bvel = oracle_species(cfg, O.baseline_cosmology(z))
    run = lpt.ICRun(ctx, lpt.ICConfig(n=n, box=box, zstart=z, lpt_order=order, lattice=lattice, plt=plt, prec=prec,
                                      seed=seed, particles=True, baryons=True, lunit=box, vunit=2.0))
    run.execute()
    p, v = run.particles()
    pb, vb = run.particles(baryons=True)
    run.close()
    q0 = species_q(n, lattice, box, 0)
    qb = species_q(n, lattice, box, 1)
    for name, got, want, q in (("dm", p, dpos, q0), ("baryon", pb, bpos, qb)):
        e, m = rel_l2(got - q, want - q), max_rel(got - q, want - q)
        print(f"  {name} Psi rel-L2 {e:.2e} max {m:.2e}", flush=True)
        assert e < TOL[prec] and m < TOL_MAX[prec], (name, e, m)
    for name, got, want in (("dm", v, dvel), ("baryon", vb, bvel)):
        e, m = rel_l2(got, want), max_rel(got, want)
        print(f"  {name} v rel-L2 {e:.2e} max {m:.2e}", flush=True)
        assert e < TOL[prec] and m < TOL_MAX[prec], (name, e, m)
    # the baryons sample the fields at other points than the dark matter
    assert rel_l2(pb - qb, p - q0) > 1e-3
    if prec == F64:
        assert np.abs(pb - bpos).max() < 1e-12 * box


@pytest.mark.gpu
@pytest.mark.parametrize("lattice,index", [(L.SC, 1), (L.BCC, 1), (L.FCC, 0), (L.FCC, 1)])
@pytest.mark.parametrize("is_pos", [1, 0])
def test_lattice_readout_index_matches_oracle(ctx, lattice, index, is_pos):
    """mfic_lattice_readout_index against the oracle's readout"""
    n, box, lunit = 16, 20.0, 3.0
    rng = np.random.default_rng(31 + lattice)
    a = O.empty_field(n)
    O.rview(a)[...] = rng.standard_normal((n, n, n))
    O.fft_r2c(a)
    h = n // 2
    k = O.kview(a)
    k[h], k[:, h], k[:, :, h] = 0, 0, 0
    O.fft_c2r(a)
    torch = pytest.importorskip("torch")
    npart = (1 << lattice) * n ** 3
    for idim in range(3):
        want = O.lattice_readout_index(a.copy(), box, lattice, index, idim, bool(is_pos), lunit)[:, idim]
        g = Grid(ctx, (n,) * 3, (box,) * 3, prec=F64)
        g.upload(a)
        dst = torch.zeros(3 * npart, dtype=torch.float64, device="cuda")  # device memory, as the entry point needs
        torch.cuda.synchronize()
        L.call("mfic_lattice_readout_index", g.h, lattice, index, idim, is_pos, lunit, C.c_void_p(dst.data_ptr()))
        ctx.sync()
        out = dst.cpu().numpy().reshape(-1, 3)
        g.close()
        err = np.max(np.abs(out[:, idim] - want)) / np.max(np.abs(want))
        assert err < 1e-12, (idim, err)


@pytest.mark.gpu
@pytest.mark.parametrize("lattice", [L.SC, L.BCC, L.FCC])
def test_baryon_particle_files(ctx, tmp_path, lattice):
    """Gadget-HDF5 (PartType0 / PartType1, per-particle Masses, zero
    MassTable) and SWIFT (PartType0 with InternalEnergy and SmoothingLength)
    of a baryon run, against the run's particles"""
    hdf = pytest.importorskip("_hdf5_reader")
    from test_output_hdf5 import reference_ids_fast
    from monofonic_amd.output import (gadget_config, gadget_units, swift_config, swift_units, write_gadget_hdf5,
                                      write_swift)
    n, box, z = 16, 40.0, 24.0
    cos = lpt.baseline_cosmology(z)
    npart = (1 << lattice) * n ** 3
    ids_dm = reference_ids_fast(n, lattice)
    for fmt in ("gadget", "swift"):
        base = lpt.ICConfig(n=n, box=box, zstart=z, lpt_order=2, lattice=lattice, prec=F32, seed=99, baryons=True)
        cfg = gadget_config(base) if fmt == "gadget" else swift_config(base, cos)
        run = lpt.ICRun(ctx, cfg, cos)
        run.execute()
        fname = str(tmp_path / f"{fmt}.hdf5")
        if fmt == "gadget":
            write_gadget_hdf5(run, fname, long_ids=True)
            _, _, munit = gadget_units(box, z)
        else:
            write_swift(run, fname)
            _, _, munit = swift_units(box, cos)
        p, v = run.particles()
        pb, vb = run.particles(baryons=True)
        run.close()
        f = hdf.File(fname)
        try:
            for grp, pp, vv, ids, om in (("PartType1", p, v, ids_dm, 0.3 - 0.045),
                                         ("PartType0", pb, vb, ids_dm + npart, 0.045)):
                c, ct = f.dataset(f"{grp}/Coordinates")
                vel, _ = f.dataset(f"{grp}/Velocities")
                got_ids, _ = f.dataset(f"{grp}/ParticleIDs")
                m, mt = f.dataset(f"{grp}/Masses")
                if fmt == "gadget":
                    assert ct == np.float32 and np.array_equal(c, pp.astype(np.float32))
                    assert np.array_equal(vel, vv.astype(np.float32))
                    assert mt == np.float32 and np.all(m == np.float32(om * munit / npart))
                else:
                    assert ct == np.float64 and np.array_equal(c, pp) and np.array_equal(vel, vv)
                    assert mt == np.float64 and np.allclose(m, om * munit / npart, rtol=1e-15, atol=0)
                assert np.array_equal(got_ids.astype(np.uint64), ids), grp
            tot = list(f.attr("Header", "NumPart_Total")[0])
            assert tot[:2] == [npart, npart]
            assert not f.attr("Header", "MassTable")[0].any()
            assert f.attr("ICs_parameters", "Baryons")[0] == 1
            if fmt == "swift":
                # output_swift.cc:86-104 with Tcmb 2.7255, YHe 0.245421, gamma 5/3
                h, ob, a = 0.7, 0.045, 1.0 / (1.0 + z)
                adec = 1.0 / (160.0 * (ob * h * h / 0.022) ** 0.4)
                tini = 2.7255 / a if a < adec else 2.7255 / a / a * adec
                mu = 4.0 / (8.0 - 5.0 * 0.245421) if tini > 1e4 else 4.0 / (1.0 + 3.0 * (1.0 - 0.245421))
                ceint = 1.3806e-16 / 1.6726e-24 * tini * 1.5 / mu / 1e10
                u, _ = f.dataset("PartType0/InternalEnergy")
                hs, _ = f.dataset("PartType0/SmoothingLength")
                assert np.allclose(u, ceint, rtol=1e-14, atol=0)
                assert np.allclose(hs, box / h / n, rtol=1e-15, atol=0)
                assert not f.exists("PartType1/InternalEnergy")
        finally:
            f.close()


@pytest.mark.gpu
def test_readout_refuses_host_memory(ctx):
    """a pageable host destination is refused before any kernel runs (a
    kernel store there would fault the device)"""
    n = 8
    g = Grid(ctx, (n,) * 3, (10.0,) * 3, prec=F64)
    host = np.zeros(3 * n ** 3)
    rc = L.load().mfic_lattice_readout_index(g.h, L.SC, 0, 0, 1, 1.0, host.ctypes.data)
    g.close()
    assert rc == -1 and b"device memory" in L.load().mfic_last_error()
