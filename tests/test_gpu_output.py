"""Particle files streamed from the device particle arrays of a run: the
Gadget-HDF5 file (output_gadget_hdf5.cc:268-318) equals the run's particles
rounded to the file's 32-bit floats, the SWIFT file (output_swift.cc:250-377)
equals them bit for bit (FP64), both with the reference's particle IDs, for
one chunk (32^3 bcc + PLT, 3LPT) and several chunks (2 x 192^3 particles)."""
import time

import numpy as np
import pytest

from monofonic_amd import BCC, F32, F64
from monofonic_amd.lpt import ICConfig, ICRun
from monofonic_amd.lpt import baseline_cosmology
from monofonic_amd.output import gadget_config, swift_config, swift_units, write_gadget_hdf5, write_swift

hdf = pytest.importorskip("_hdf5_reader")
from test_output_hdf5 import reference_ids_fast  # noqa: E402


@pytest.mark.gpu
@pytest.mark.parametrize("n,order,plt,prec,long_ids", [(32, 3, True, F32, False), (32, 2, False, F64, True),
                                                       (192, 2, False, F32, True)])
def test_gadget_file_from_device(ctx, tmp_path, n, order, plt, prec, long_ids):
    print(f"Gadget-HDF5 from the device: n={n} order={order} plt={plt} prec={prec}", flush=True)
    cfg = gadget_config(ICConfig(n=n, box=2.0 * n, zstart=24.0, lpt_order=order, lattice=BCC, plt=plt, prec=prec,
                                 seed=555))
    run = ICRun(ctx, cfg)
    run.execute()
    fname = str(tmp_path / "ics.hdf5")
    t = time.perf_counter()
    write_gadget_hdf5(run, fname, long_ids=long_ids)
    dt = time.perf_counter() - t
    pos, vel = run.particles()
    run.close()
    npart = 2 * n ** 3
    print(f"  wrote {npart} particles in {dt:.3f} s", flush=True)
    f = hdf.File(fname)
    try:
        c, ct = f.dataset("PartType1/Coordinates")
        v, _ = f.dataset("PartType1/Velocities")
        ids, it = f.dataset("PartType1/ParticleIDs")
        assert ct == np.float32 and c.shape == (npart, 3)
        assert np.array_equal(c, pos.astype(np.float32))
        assert np.array_equal(v, vel.astype(np.float32))
        assert it == (np.uint64 if long_ids else np.uint32)
        assert np.array_equal(ids.astype(np.uint64), reference_ids_fast(n, 1))
        assert list(f.attr("Header", "NumPart_Total")[0]) == [0, npart, 0, 0, 0, 0]
        assert f.attr("Header", "BoxSize")[0][0] == 2.0 * n
    finally:
        f.close()


@pytest.mark.gpu
@pytest.mark.parametrize("n,order,plt,prec,long_ids", [(32, 3, True, F32, True), (192, 2, False, F64, False)])
def test_swift_file_from_device(ctx, tmp_path, n, order, plt, prec, long_ids):
    print(f"SWIFT from the device: n={n} order={order} plt={plt} prec={prec}", flush=True)
    cos = baseline_cosmology(24.0)
    cfg = swift_config(ICConfig(n=n, box=2.0 * n, zstart=24.0, lpt_order=order, lattice=BCC, plt=plt, prec=prec,
                                seed=556), cos)
    run = ICRun(ctx, cfg, cos)
    run.execute()
    fname = str(tmp_path / "ics_swift.hdf5")
    t = time.perf_counter()
    write_swift(run, fname, long_ids=long_ids)
    dt = time.perf_counter() - t
    pos, vel = run.particles()
    run.close()
    npart = 2 * n ** 3
    print(f"  wrote {npart} particles in {dt:.3f} s", flush=True)
    _, _, munit = swift_units(2.0 * n, cos)
    f = hdf.File(fname)
    try:
        c, ct = f.dataset("PartType1/Coordinates")
        v, _ = f.dataset("PartType1/Velocities")
        ids, it = f.dataset("PartType1/ParticleIDs")
        m, _ = f.dataset("PartType1/Masses")
        assert ct == np.float64 and np.array_equal(c, pos) and np.array_equal(v, vel)
        assert it == (np.uint64 if long_ids else np.uint32)
        assert np.array_equal(ids.astype(np.uint64), reference_ids_fast(n, 1))
        assert np.all(m == 0.3 * munit / npart)
        assert list(f.attr("Header", "NumPart_ThisFile")[0]) == [0, npart, 0, 0, 0, 0, 0]
    finally:
        f.close()
