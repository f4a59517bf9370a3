"""Tabulated transfer function on the device (the CAMB_file plugin,
transfer_CAMB_file.cc) against the CPU FP64 oracle: phi1, whole 2LPT runs,
and the error for a table that stops short of the grid's largest |k|."""
import numpy as np
import pytest

from monofonic_amd import F32, F64, Grid, _lib as L, lpt, transfer
from oracle import oracle as O

from test_oracle_transfer import BASE, wiggly_table


def rel_l2(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


TOL = {F64: 1e-10, F32: 1e-5}


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [F64, F32])
def test_phi1_table_matches_oracle(ctx, prec):
    n, box, seed = 32, 100.0, 12345
    k, T = wiggly_table()
    oc = O.cosmology(transfer=(k, T), **dict(BASE, zstart=49.0))
    cfg = O.ic_cfg(n, box, 1, seed)
    phi, _, _, _ = O.potentials(cfg, oc, O.ngenic_fill(seed, n))
    coef = O.lpt_coefs(cfg, oc)
    g = Grid(ctx, (n,) * 3, (box,) * 3, prec=prec)
    try:
        lpt.ngenic_fill(g, seed)
        lpt.phi1_from_noise(g, g, transfer.with_table(lpt.baseline_cosmology(49.0), k, T))
        err = rel_l2(g.download_k() * coef.g1, O.kview(phi))
        print(f"phi1 with a tabulated T(k), prec {prec}: rel-L2 {err:.2e}")
        assert err < TOL[prec]
    finally:
        g.close()


@pytest.mark.gpu
def test_2lpt_fields_table_match_oracle(ctx):
    n, box, z, seed = 32, 100.0, 24.0, 4242
    k, T = wiggly_table()
    oc = O.cosmology(transfer=(k, T), **BASE)
    cfg = O.ic_cfg(n, box, 2, seed)
    fields, _, _, _ = O.ic_run(cfg, oc)
    cosmo = transfer.with_table(lpt.baseline_cosmology(z), k, T)
    run = lpt.ICRun(ctx, lpt.ICConfig(n=n, box=box, zstart=z, lpt_order=2, prec=F64, seed=seed), cosmo)
    try:
        run.execute()
        for c in range(6):
            err = rel_l2(run.field(c).download_r(), O.rview(fields[c]))
            assert err < 50 * TOL[F64], (c, err)
    finally:
        run.close()


@pytest.mark.gpu
def test_table_must_cover_the_grid(ctx):
    n, box = 32, 100.0                       # largest |k| = sqrt(3) pi n / box = 1.74 h/Mpc
    k, T = wiggly_table(kmax=1.0)
    cosmo = transfer.with_table(lpt.baseline_cosmology(24.0), k, T)
    with pytest.raises(L.MficError) as e:
        lpt.ICRun(ctx, lpt.ICConfig(n=n, box=box, zstart=24.0, lpt_order=1, prec=F64), cosmo)
    assert e.value.code == -6
    g = Grid(ctx, (n,) * 3, (box,) * 3, prec=F64)
    try:
        lpt.ngenic_fill(g, 1)
        with pytest.raises(L.MficError):
            lpt.phi1_from_noise(g, g, cosmo)
    finally:
        g.close()
