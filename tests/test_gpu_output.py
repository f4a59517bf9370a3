"""Gadget-HDF5 output streamed from the device particle arrays of a run
(output_gadget_hdf5.cc:268-318): the file equals the run's particles
rounded to the file's 32-bit floats, with the reference's particle IDs, for
one chunk (32^3 bcc + PLT, 3LPT) and several chunks (2 x 192^3 particles)."""
import time

import numpy as np
import pytest

from monofonic_amd import BCC, F32, F64
from monofonic_amd.lpt import ICConfig, ICRun
from monofonic_amd.output import gadget_config, write_gadget_hdf5

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
