"""Helpers shared by the GPU parity tests (not a test module).

Sampled reads of device buffers go through the HIP runtime's hipMemcpy, so a
512^3 run can be compared at a few thousand particles or cells without
copying its 13 GB of particle arrays to the host.
"""
import ctypes as C
import os

import numpy as np

from monofonic_amd import _lib as L
from monofonic_amd.lpt import _hip

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def max_rel(a, b):
    """max |a - b| over max |b|"""
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def load_sample(name):
    with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def read_rows(ctx, dptr, rows, width, dtype=np.float64):
    """rows[i]-th record of `width` values of a device array at dptr"""
    hip = _hip()
    ctx.sync()
    item = np.dtype(dtype).itemsize * width
    out = np.empty((len(rows), width), dtype=dtype)
    for i, r in enumerate(np.asarray(rows, dtype=np.int64)):
        rc = hip.hipMemcpy(C.c_void_p(out[i].ctypes.data), C.c_void_p(dptr + int(r) * item), C.c_size_t(item), 2)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy failed ({rc})")
    return out


SUBLATTICES = {L.SC: [(0, 0, 0)], L.BCC: [(0, 0, 0), (0.5, 0.5, 0.5)],
               L.FCC: [(0, 0, 0), (0, 0.5, 0.5), (0.5, 0, 0.5), (0.5, 0.5, 0)]}


def lattice_q(n, lattice, particles, lunit=1.0):
    """Lagrangian positions ((ijk + shift) / n) * lunit of particle indices
    (particle_generator.hh:47-63, 499-544; sub-lattice s holds indices
    s n^3 .. (s+1) n^3 - 1 at the lattice_shifts offset s)"""
    p = np.asarray(particles, dtype=np.int64)
    s = p // n ** 3
    c = p % n ** 3
    ijk = np.stack(np.unravel_index(c, (n, n, n)), axis=1).astype(np.float64)
    shift = np.array(SUBLATTICES[lattice], dtype=np.float64)[s]
    return (ijk + shift) / n * lunit


def grid_device_ptr(g):
    p, nr = C.c_void_p(), C.c_size_t()
    L.call("mfic_grid_device_ptr", g.h, C.byref(p), C.byref(nr))
    return p.value, nr.value


def sample_k(g, modes):
    """complex values of a k-space grid at (i, j, k) modes"""
    k = g.download_k()
    return k[modes[:, 0], modes[:, 1], modes[:, 2]].astype(np.complex128)


def sample_r(g, cells):
    r = g.download_r()
    return r[cells[:, 0], cells[:, 1], cells[:, 2]].astype(np.float64)
