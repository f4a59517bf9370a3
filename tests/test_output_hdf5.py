"""The Gadget-HDF5 writer (output_gadget_hdf5.cc:36-318 with HDF_IO.hh) from
host arrays, read back through the HDF5 C library: datasets, their file
datatypes, the particle ID rule (particle_generator.hh:229-251), the header
and the run description."""
import math
import os

import numpy as np
import pytest

from monofonic_amd import _lib as L
from monofonic_amd import lpt
from monofonic_amd.lpt import ICConfig
from monofonic_amd.output import gadget_config, gadget_units, write_gadget_hdf5_host
from oracle import oracle as O

hdf = pytest.importorskip("_hdf5_reader")
try:
    hdf.lib()
except OSError:
    pytest.skip("HDF5 C library absent", allow_module_level=True)


def reference_ids(n, lattice):
    """the reference's loop (particle_generator.hh:233-251) with IDoffset 0:
    particle ipcount + iload * n^3 of cell (i, j, k) gets
    overload * get_cell_idx_1d(i, j, k) + iload"""
    overload = 1 << lattice
    ids = np.zeros(overload * n ** 3, dtype=np.uint64)
    ipcount = 0
    for i in range(n):
        for j in range(n):
            for k in range(n):
                cell = (i * n + j) * n + k
                for iload in range(overload):
                    ids[ipcount + iload * n ** 3] = overload * cell + iload
                ipcount += 1
    return ids


def reference_ids_fast(n, lattice):
    """the same rule vectorised: index iload * n^3 + cell -> overload * cell + iload"""
    overload = 1 << lattice
    cell = np.arange(n ** 3, dtype=np.uint64)
    return np.concatenate([overload * cell + iload for iload in range(overload)])


@pytest.mark.parametrize("lattice", [0, 1])
def test_vectorised_id_rule_equals_the_loop(lattice):
    assert np.array_equal(reference_ids_fast(6, lattice), reference_ids(6, lattice))


def test_gadget_units():
    lu, vu, mu = gadget_units(100.0, 49.0)
    a = 1.0 / 50.0
    assert lu == 100.0 and vu == 100.0 / math.sqrt(a) and mu == 27.7519737 * 100.0 ** 3


@pytest.mark.parametrize("lattice,long_ids,compat", [(0, False, False), (1, True, False), (1, False, True)])
def test_gadget_file_from_host_arrays(tmp_path, lattice, long_ids, compat):
    n, box, z = 8, 100.0, 24.0
    cfg = gadget_config(ICConfig(n=n, box=box, zstart=z, lpt_order=2, lattice=lattice, seed=7))
    npart = (1 << lattice) * n ** 3
    rng = np.random.default_rng(lattice)
    pos = rng.uniform(0, box, (npart, 3))
    vel = rng.normal(0, 300.0, (npart, 3))
    fname = str(tmp_path / "ics.hdf5")
    cos = lpt.baseline_cosmology(z)
    write_gadget_hdf5_host(fname, cfg, cos, pos, vel, long_ids=long_ids, gadget2_compat=compat)
    f = hdf.File(fname)
    try:
        c, dt = f.dataset("PartType1/Coordinates")
        assert dt == np.float32 and c.shape == (npart, 3)
        assert np.array_equal(c, pos.astype(np.float32))
        v, dt = f.dataset("PartType1/Velocities")
        assert dt == np.float32 and np.array_equal(v, vel.astype(np.float32))
        ids, dt = f.dataset("PartType1/ParticleIDs")
        assert dt == (np.uint64 if long_ids else np.uint32)
        assert np.array_equal(ids.astype(np.uint64), reference_ids(n, lattice))
        assert not f.exists("PartType1/Masses")
        # header: 1-element arrays except the 6-arrays (output_gadget_hdf5.cc:175-210)
        np_this, dt, shape = f.attr("Header", "NumPart_ThisFile")
        assert shape == (6,) and dt == (np.uint32 if compat else np.uint64)
        assert list(np_this) == [0, npart, 0, 0, 0, 0]
        assert list(f.attr("Header", "NumPart_Total")[0]) == [0, npart, 0, 0, 0, 0]
        if compat:
            assert list(f.attr("Header", "NumPart_Total_HighWord")[0]) == [0] * 6
        oc = O.baseline_cosmology(z)
        mass, dt, _ = f.attr("Header", "MassTable")
        _, _, munit = gadget_units(box, z)
        assert dt == np.float64 and mass[1] == 0.3 * munit / npart and mass[0] == 0 and mass[2] == 0
        a = 1.0 / (1.0 + z)
        for name, want in (("Time", a), ("Redshift", 1.0 / a - 1.0), ("BoxSize", box), ("Omega0", 0.3),
                           ("OmegaLambda", oc.Omega_DE), ("HubbleParam", 0.7)):
            val, dt, shape = f.attr("Header", name)
            assert dt == np.float64 and shape == (1,) and val[0] == want, name
        for name, want in (("Flag_Sfr", 0), ("Flag_Feedback", 0), ("Flag_Cooling", 0), ("NumFilesPerSnapshot", 1),
                           ("Flag_StellarAge", 0), ("Flag_Metals", 0), ("Flag_Entropy_ICs", 0)):
            val, dt, shape = f.attr("Header", name)
            assert dt == np.int32 and shape == (1,) and val[0] == want, name
        # run description (scalar attributes, output_gadget_hdf5.cc:140-170)
        assert f.attr("ICs_parameters", "LPT Order")[:2] == (2, np.dtype(np.int32))
        assert f.attr("ICs_parameters", "Grid Resolution")[0] == n
        assert f.attr("ICs_parameters", "Particle Load")[0] == (b"bcc" if lattice else b"sc")
        assert f.attr("ICs_parameters", "Random Generator")[0] == b"NGENIC"
        assert f.attr("ICs_parameters", "Mode Fixing")[0] == 0
    finally:
        f.close()


def test_writer_rejects_other_units(tmp_path):
    cfg = ICConfig(n=4, box=50.0, zstart=24.0, particles=True)  # lunit = vunit = 1: not the plugin's units
    pos = np.zeros((64, 3))
    with pytest.raises(L.MficError) as e:
        write_gadget_hdf5_host(str(tmp_path / "x.hdf5"), cfg, lpt.baseline_cosmology(24.0), pos, pos)
    assert e.value.code == -6
    assert not os.path.exists(tmp_path / "x.hdf5")


# ------------------------------------------------------------------ SWIFT --
def test_swift_units():
    from monofonic_amd.output import swift_units
    cos = lpt.baseline_cosmology(24.0)
    lu, vu, mu = swift_units(100.0, cos)
    assert lu == 100.0 / 0.7 and vu == 100.0 and mu == 27.7536609198 * 100.0 ** 3 / 0.7


@pytest.mark.parametrize("lattice,long_ids", [(0, True), (1, False)])
def test_swift_file_from_host_arrays(tmp_path, lattice, long_ids):
    """the layout of output_swift.cc:36-377 with HDF_IO.hh: Units, FP64
    vectors, per-particle Masses, chunked datasets with Fletcher-32 (filter
    id 3), the 7-entry header"""
    from monofonic_amd.output import swift_config, swift_units, write_swift_host
    n, box, z = 8, 100.0, 24.0
    cos = lpt.baseline_cosmology(z)
    cfg = swift_config(ICConfig(n=n, box=box, zstart=z, lpt_order=2, lattice=lattice, seed=7), cos)
    npart = (1 << lattice) * n ** 3
    rng = np.random.default_rng(5 + lattice)
    pos = rng.uniform(0, box / 0.7, (npart, 3))
    vel = rng.normal(0, 300.0, (npart, 3))
    fname = str(tmp_path / "ics_swift.hdf5")
    write_swift_host(fname, cfg, cos, pos, vel, long_ids=long_ids)
    _, _, munit = swift_units(box, cos)
    f = hdf.File(fname)
    try:
        c, dt = f.dataset("PartType1/Coordinates")
        assert dt == np.float64 and np.array_equal(c, pos)
        v, dt = f.dataset("PartType1/Velocities")
        assert dt == np.float64 and np.array_equal(v, vel)
        ids, dt = f.dataset("PartType1/ParticleIDs")
        assert dt == (np.uint64 if long_ids else np.uint32)
        assert np.array_equal(ids.astype(np.uint64), reference_ids(n, lattice))
        m, dt = f.dataset("PartType1/Masses")
        assert dt == np.float64 and m.shape == (npart,) and np.all(m == 0.3 * munit / npart)
        # chunks: 256 KB of vector rows, 1 MB of scalars, cut to the length
        assert f.storage("PartType1/Coordinates") == ((min(npart, 32768), 3), [3])
        assert f.storage("PartType1/Masses") == ((min(npart, 131072),), [3])
        assert f.storage("PartType1/ParticleIDs") == ((min(npart, (1 << 20) // (8 if long_ids else 4)),), [3])
        for name, want in (("Unit mass in cgs (U_M)", 1.98841e43), ("Unit length in cgs (U_L)", 3.08567758149e24),
                           ("Unit time in cgs (U_t)", 3.08567758149e19), ("Unit current in cgs (U_I)", 1.0),
                           ("Unit temperature in cgs (U_T)", 1.0)):
            val, dt, shape = f.attr("Units", name)
            assert dt == np.float64 and shape == () and val == want, name
        assert f.attr("Header", "Dimension")[:3] == (3, np.dtype(np.int32), ())
        assert f.attr("Header", "BoxSize")[0] == box / 0.7
        tot, dt, shape = f.attr("Header", "NumPart_Total")
        assert dt == np.uint32 and shape == (7,) and list(tot) == [0, npart, 0, 0, 0, 0, 0]
        assert list(f.attr("Header", "NumPart_Total_HighWord")[0]) == [0] * 7
        this, dt, _ = f.attr("Header", "NumPart_ThisFile")
        assert dt == np.uint64 and list(this) == [0, npart, 0, 0, 0, 0, 0]
        mass, dt, shape = f.attr("Header", "MassTable")
        assert dt == np.float64 and shape == (7,) and not mass.any()
        for name, want, dtype in (("Time", 1.0 / (1.0 + z), np.float64), ("Redshift", z, np.float64),
                                  ("Flag_Entropy_ICs", 0, np.int32), ("NumFilesPerSnapshot", 1, np.int32)):
            val, dt, shape = f.attr("Header", name)
            assert dt == dtype and shape == (1,) and val[0] == want, name
        assert f.attr("ICs_parameters", "Particle Load")[0] == (b"bcc" if lattice else b"sc")
    finally:
        f.close()


def test_swift_writer_rejects_gadget_units(tmp_path):
    from monofonic_amd.output import write_swift_host
    cfg = gadget_config(ICConfig(n=4, box=50.0, zstart=24.0))
    pos = np.zeros((64, 3))
    with pytest.raises(L.MficError) as e:
        write_swift_host(str(tmp_path / "y.hdf5"), cfg, lpt.baseline_cosmology(24.0), pos, pos)
    assert e.value.code == -6
