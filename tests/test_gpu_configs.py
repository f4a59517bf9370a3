"""The benchmark configurations end to end on the device against the CPU FP64
oracle.

C1  64^3 2LPT, box 100 Mpc/h, z_start 49, seed 12345, sc particles: the
    whole run against the oracle's run (FP64 and FP32).
C2  256^3 3LPT sc, box 500 Mpc/h, z_start 24: the six potentials, and the
    displacement and velocity of each LPT order separately (Psi1, Psi2, Psi3
    with the A3 curl) through mfic_lpt_component, against a sampled oracle
    run (tests/golden/c2_sample.npz).
C3  512^3 3LPT + PLT bcc (2 x 512^3 particles), box 1000 Mpc/h, z_start 24:
    FP32 bulk with FP64 q+Psi (the benchmarked path) and FP64, particles and
    potentials against a sampled oracle run (tests/golden/c3_sample.npz).
C4  three boxes of the 48-box 2LPT FP32 workload (one per size) against the
    oracle, and the full 48-box HIP graph bit-identical to the eager run.
C5  (tests/test_gpu_adjoint.py) 256^3 2LPT adjoint against FP64 central
    differences.

Tolerances: FP64 rel-L2 1e-10 (different FFT algorithms and summation order);
FP32 rel-L2 1e-5 on displacements and velocities (the FP32 target), max
error relative to the largest value 1e-4.
"""
import ctypes as C

import numpy as np
import pytest

from monofonic_amd import F32, F64, _lib as L, lpt
from oracle import oracle as O

from _gpu_helpers import lattice_q, load_sample, max_rel, read_rows, rel_l2, sample_k, sample_r

TOL = {F64: 1e-10, F32: 1e-5}
TOL_MAX = {F64: 1e-9, F32: 1e-4}


def pot_tol(prec, w):
    """second- and third-order potentials in FP64: 5e-9 (their sources are
    sums of products that cancel to a few digits)"""
    return 5e-9 if prec == F64 and w > 0 else TOL[prec]


def _flush(msg):
    print(msg, flush=True)


# -------------------------------------------------------------------- C1 --
@pytest.mark.gpu
@pytest.mark.parametrize("prec", [F64, F32])
def test_c1_particles_match_oracle(ctx, prec):
    n, box, z, seed = 64, 100.0, 49.0, 12345
    _flush(f"C1 64^3 2LPT sc particles prec={prec}")
    cfg = O.ic_cfg(n, box, 2, seed)
    _, pos, vel, _ = O.ic_run(cfg, O.baseline_cosmology(z), particles=True)
    run = lpt.ICRun(ctx, lpt.ICConfig(n=n, box=box, zstart=z, lpt_order=2, prec=prec, seed=seed, particles=True))
    run.execute()
    p, v = run.particles()
    run.close()
    pos, vel = pos.reshape(-1, 3), vel.reshape(-1, 3)
    q = lattice_q(n, L.SC, np.arange(n ** 3))
    e_psi, e_v = rel_l2(p - q, pos - q), rel_l2(v, vel)
    m_psi, m_v = max_rel(p - q, pos - q), max_rel(v, vel)
    _flush(f"C1 prec {prec}: Psi rel-L2 {e_psi:.2e} max {m_psi:.2e}; v rel-L2 {e_v:.2e} max {m_v:.2e}")
    assert e_psi < TOL[prec] and e_v < TOL[prec]
    assert m_psi < TOL_MAX[prec] and m_v < TOL_MAX[prec]
    if prec == F64:
        assert np.abs(p - pos).max() < 1e-12


# -------------------------------------------------------------------- C2 --
def _order_coef(coef, order):
    """LPT coefficients with every other order's growth factor zeroed: the
    assembly then returns one order's displacement / velocity"""
    k = L.LptCoef()
    for name, _ in L.LptCoef._fields_:
        setattr(k, name, getattr(coef, name))
    if order != 1:
        k.g1 = 0.0
    if order != 2:
        k.g2 = 0.0
    if order != 3:
        k.g3 = k.g3c = 0.0
    return k


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [F64, F32])
def test_c2_potentials_and_orders_match_oracle(ctx, prec):
    s = load_sample("c2_sample.npz")
    n, box, z, seed = int(s["n"]), float(s["box"]), float(s["zstart"]), int(s["seed"])
    _flush(f"C2 {n}^3 3LPT prec={prec}")
    cos = lpt.baseline_cosmology(z)
    coef = lpt.lpt_coefficients(cos, n, box, 3)
    run = lpt.ICRun(ctx, lpt.ICConfig(n=n, box=box, zstart=z, lpt_order=3, prec=prec, seed=seed), cos)
    run.execute()
    scale = [coef.g1, coef.g2, coef.g3, coef.g3c, coef.g3c, coef.g3c]
    for w in range(6):
        got = sample_k(run.potential(w), s["modes"]) * scale[w]
        err = rel_l2(got, s["potentials"][w])
        _flush(f"C2 prec {prec} potential {w}: rel-L2 {err:.2e}")
        assert err < pot_tol(prec, w), (w, err)
    pots = [run.potential(w) for w in range(6)]
    out = lpt.Grid(ctx, (n,) * 3, (box,) * 3, prec=prec)
    total = np.zeros((6, s["cells"].shape[0]))
    for order in (1, 2, 3):
        k = _order_coef(coef, order)
        for comp in range(6):
            lpt.lpt_component(ctx, 3, comp, k, (pots[0], pots[1], pots[2], pots[3:]), out, 1.0 / box)
            out.FourierTransformBackward()
            got = sample_r(out, s["cells"])
            want = s["fields"][order - 1, comp]
            err, mx = rel_l2(got, want), max_rel(got, want)
            _flush(f"C2 prec {prec} order {order} {'psi' if comp < 3 else 'v'}{comp % 3}: rel-L2 {err:.2e} max {mx:.2e}")
            assert err < TOL[prec] and mx < TOL_MAX[prec], (order, comp, err, mx)
            total[comp] += want
    for comp in range(6):
        err = rel_l2(sample_r(run.field(comp), s["cells"]), total[comp])
        assert err < TOL[prec], (comp, err)
    out.close()
    run.close()


# -------------------------------------------------------------------- C3 --
def _set_plans(ctx, n, prec, variant):
    """every strided pass of the n^3 run (N-grid pencils, shifted
    sub-lattice pass, padded convolver passes) on radix plan `variant`, and
    the fused source row passes in form `variant` (1: one half-length
    transform per row instead of two rows per full-length transform; -1:
    back to the defaults)"""
    from monofonic_amd import batch
    for length in (n, 3 * n // 2):
        if batch.plan_count(length) > 1:
            for p in L.PASSES:
                batch.set_plan(ctx, length, p, prec, variant)


@pytest.mark.gpu
@pytest.mark.parametrize("plans", ["default", "alternative"])
@pytest.mark.parametrize("prec", [F32, F64])
def test_c3_particles_match_oracle(ctx, prec, plans):
    s = load_sample("c3_sample.npz")
    n, box, z, seed = int(s["n"]), float(s["box"]), float(s["zstart"]), int(s["seed"])
    _flush(f"C3 {n}^3 3LPT+PLT bcc particles prec={prec} radix plans {plans}")
    run = lpt.ICRun(ctx, lpt.ICConfig(n=n, box=box, zstart=z, lpt_order=3, lattice=L.BCC, plt=True, prec=prec,
                                      seed=seed, particles=True))
    if plans == "alternative":
        _set_plans(ctx, n, prec, 1)
    try:
        run.execute()
    finally:
        _set_plans(ctx, n, prec, -1)
    pos_d, vel_d, npart = run.particles_device()
    assert npart == 2 * n ** 3
    parts = s["particles"]
    p = read_rows(ctx, pos_d, parts, 3)
    v = read_rows(ctx, vel_d, parts, 3)
    q = lattice_q(n, L.BCC, parts, float(s["lunit"]))
    e_psi, m_psi = rel_l2(p - q, s["pos"] - q), max_rel(p - q, s["pos"] - q)
    e_v, m_v = rel_l2(v, s["vel"]), max_rel(v, s["vel"])
    _flush(f"C3 prec {prec}: Psi rel-L2 {e_psi:.2e} max {m_psi:.2e}; v rel-L2 {e_v:.2e} max {m_v:.2e} "
           f"({len(parts)} particles)")
    assert e_psi < TOL[prec] and e_v < TOL[prec]
    assert m_psi < TOL_MAX[prec] and m_v < TOL_MAX[prec]
    cos = lpt.baseline_cosmology(z)
    coef = lpt.lpt_coefficients(cos, n, box, 3)
    scale = [coef.g1, coef.g2, coef.g3, coef.g3c, coef.g3c, coef.g3c]
    for w in range(6):
        err = rel_l2(sample_k(run.potential(w), s["modes"]) * scale[w], s["potentials"][w])
        _flush(f"C3 prec {prec} potential {w}: rel-L2 {err:.2e}")
        assert err < pot_tol(prec, w), (w, err)
    run.close()


# -------------------------------------------------------------------- C4 --
@pytest.mark.gpu
def test_c4_one_box_per_size_matches_oracle(ctx):
    s = load_sample("c4_sample.npz")
    for slot in range(3):
        i, n, seed = (int(x) for x in s[f"box{slot}"])
        box = float(s[f"L{slot}"])
        _flush(f"C4 box {i}: {n}^3 L={box:.3f} seed {seed} FP32")
        run = lpt.ICRun(ctx, lpt.ICConfig(n=n, box=box, zstart=24.0, lpt_order=2, prec=F32, seed=seed))
        run.execute()
        for comp in range(6):
            got = sample_r(run.field(comp), s[f"cells{slot}"])
            want = s[f"fields{slot}"][comp]
            err, mx = rel_l2(got, want), max_rel(got, want)
            assert err < TOL[F32] and mx < TOL_MAX[F32], (i, comp, err, mx)
        _flush(f"C4 box {i}: fields within rel-L2 {TOL[F32]:g}")
        run.close()


@pytest.mark.gpu
def test_c4_full_graph_bit_identical_to_eager(ctx):
    """all 48 boxes (415,236,096 cells) replayed from one HIP graph equal the
    eager run bit for bit, compared on the device"""
    torch = pytest.importorskip("torch")
    from monofonic_amd.batch import Batch, batch_boxes
    from _gpu_helpers import grid_device_ptr

    class _Dev:
        def __init__(self, ptr, count):
            self.__cuda_array_interface__ = {"shape": (count,), "typestr": "<i4", "data": (ptr, False),
                                             "version": 3}

    _flush("C4 48-box graph vs eager")
    b = Batch(ctx, batch_boxes(), lpt_order=2, prec=F32, tune=False)
    try:
        b.execute()
        views, ref = [], []
        for r in b.runs:
            for c in range(6):
                ptr, nr = grid_device_ptr(r.field(c))
                t = torch.as_tensor(_Dev(ptr, nr), device="cuda")
                views.append(t)
                ref.append(t.clone())
        torch.cuda.synchronize()
        b.capture()
        for rep in range(2):
            for t in views:
                t.zero_()
            torch.cuda.synchronize()
            b.launch()
            bad = [k for k, (t, r) in enumerate(zip(views, ref)) if not torch.equal(t, r)]
            assert not bad, (rep, bad[:5])
            _flush(f"C4 graph replay {rep}: 48 boxes x 6 fields bit-identical")
    finally:
        b.close()


# ------------------------------------------------------ 384^3 (576 padded) --
@pytest.mark.gpu
@pytest.mark.parametrize("prec", [F64, F32])
def test_384_potentials_match_oracle(ctx, prec):
    """the padded row length whose transforms take 32 lanes (576^3 padded),
    against a sampled oracle run (tests/golden/p384_sample.npz)"""
    s = load_sample("p384_sample.npz")
    n, box, z, seed = int(s["n"]), float(s["box"]), float(s["zstart"]), int(s["seed"])
    _flush(f"384^3 3LPT potentials prec={prec}")
    cos = lpt.baseline_cosmology(z)
    coef = lpt.lpt_coefficients(cos, n, box, 3)
    run = lpt.ICRun(ctx, lpt.ICConfig(n=n, box=box, zstart=z, lpt_order=3, prec=prec, seed=seed), cos)
    run.execute()
    scale = [coef.g1, coef.g2, coef.g3, coef.g3c, coef.g3c, coef.g3c]
    for w in range(6):
        err = rel_l2(sample_k(run.potential(w), s["modes"]) * scale[w], s["potentials"][w])
        _flush(f"384^3 prec {prec} potential {w}: rel-L2 {err:.2e}")
        assert err < pot_tol(prec, w), (w, err)
    run.close()
