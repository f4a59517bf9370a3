"""Particle output on top of the C ABI: the Gadget-HDF5 writer of the
gadget_hdf5 / gadget_hdf5_double output plugins (output_gadget_hdf5.cc:36-318)
and the SWIFT writer (output_swift.cc:36-377).

``gadget_units(box, zstart)`` gives the Gadget plugin's position / velocity /
mass units (output_gadget_hdf5.cc:79-85), ``swift_units(box, cosmo)`` the
SWIFT plugin's (output_swift.cc:63-69); a run meant for a writer is created
with ``lunit`` and ``vunit`` set to the first two (``gadget_config`` /
``swift_config`` do it).  ``write_gadget_hdf5(run, fname)`` and
``write_swift(run, fname)`` stream the run's device particles into the file;
the ``*_host`` forms write host arrays.  The HDF5 C library is loaded by
libmfic at run time; the writing happens there, not in Python.
"""
from __future__ import annotations

import ctypes as C
import dataclasses

import numpy as np

from . import _lib as L
from .lpt import ICConfig, ICRun


def gadget_units(box: float, zstart: float):
    lu, vu, mu = C.c_double(), C.c_double(), C.c_double()
    L.call("mfic_gadget_units", box, zstart, C.byref(lu), C.byref(vu), C.byref(mu))
    return lu.value, vu.value, mu.value


def gadget_config(cfg: ICConfig) -> ICConfig:
    """the same run configuration with the writer's output units and particle
    output switched on"""
    lu, vu, _ = gadget_units(cfg.box, cfg.zstart)
    return dataclasses.replace(cfg, particles=True, lunit=lu, vunit=vu)


def _opts(double_precision: bool, long_ids: bool, gadget2_compat: bool) -> L.GadgetOpts:
    o = L.GadgetOpts()
    o.double_precision, o.long_ids, o.gadget2_compat = int(double_precision), int(long_ids), int(gadget2_compat)
    return o


def write_gadget_hdf5(run: ICRun, fname: str, double_precision: bool = False, long_ids: bool = False,
                      gadget2_compat: bool = False) -> None:
    """write_particle_data + header of an executed particle run"""
    o = _opts(double_precision, long_ids, gadget2_compat)
    L.call("mfic_gadget_write", run.h, fname.encode(), C.byref(o))


def write_gadget_hdf5_host(fname: str, cfg: ICConfig, cosmo: L.CosmoParams, pos: np.ndarray, vel: np.ndarray,
                           double_precision: bool = False, long_ids: bool = False,
                           gadget2_compat: bool = False) -> None:
    """the same file from host FP64 (npart, 3) arrays of cfg's lattice"""
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    vel = np.ascontiguousarray(vel, dtype=np.float64)
    npart = (1 << cfg.lattice) * cfg.n ** 3
    if pos.size != 3 * npart or vel.size != 3 * npart:
        raise ValueError(f"expected {npart} x 3 positions and velocities")
    o = _opts(double_precision, long_ids, gadget2_compat)
    c = cfg.to_c()
    cosmo.zstart = cfg.zstart
    L.call("mfic_gadget_write_host", fname.encode(), C.byref(c), C.byref(cosmo), C.byref(o), pos.ctypes.data,
           vel.ctypes.data)


def swift_units(box: float, cosmo: L.CosmoParams):
    lu, vu, mu = C.c_double(), C.c_double(), C.c_double()
    L.call("mfic_swift_units", box, C.byref(cosmo), C.byref(lu), C.byref(vu), C.byref(mu))
    return lu.value, vu.value, mu.value


def swift_config(cfg: ICConfig, cosmo: L.CosmoParams) -> ICConfig:
    """the same run configuration in the SWIFT writer's units, particles on"""
    lu, vu, _ = swift_units(cfg.box, cosmo)
    return dataclasses.replace(cfg, particles=True, lunit=lu, vunit=vu)


def write_swift(run: ICRun, fname: str, long_ids: bool = True) -> None:
    """the SWIFT initial-condition file of an executed particle run
    (UseLongids defaults to true in that plugin)"""
    o = L.SwiftOpts()
    o.long_ids = int(long_ids)
    L.call("mfic_swift_write", run.h, fname.encode(), C.byref(o))


def write_swift_host(fname: str, cfg: ICConfig, cosmo: L.CosmoParams, pos: np.ndarray, vel: np.ndarray,
                     long_ids: bool = True) -> None:
    """the same file from host FP64 (npart, 3) arrays of cfg's lattice"""
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    vel = np.ascontiguousarray(vel, dtype=np.float64)
    npart = (1 << cfg.lattice) * cfg.n ** 3
    if pos.size != 3 * npart or vel.size != 3 * npart:
        raise ValueError(f"expected {npart} x 3 positions and velocities")
    o = L.SwiftOpts()
    o.long_ids = int(long_ids)
    c = cfg.to_c()
    cosmo.zstart = cfg.zstart
    L.call("mfic_swift_write_host", fname.encode(), C.byref(c), C.byref(cosmo), C.byref(o), pos.ctypes.data,
           vel.ctypes.data)
