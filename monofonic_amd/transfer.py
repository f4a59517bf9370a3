"""Tabulated transfer functions: the CAMB_file plugin of the reference
(transfer_CAMB_file.cc).  ``read_camb_file`` reads the table the plugin reads;
``with_table`` attaches (k, T) to the cosmology parameters, after which every
call that takes them (phi1, runs, the 2LPT map) evaluates T(k) as the
natural cubic spline of ln T over ln k on the device instead of Eisenstein &
Hu.
"""
from __future__ import annotations

import ctypes as C
from typing import Tuple

import numpy as np

from . import _lib as L

# columns of a CAMB transfer file (transfer_CAMB_file.cc:62-75)
CAMB_COLUMNS = ("k", "cdm", "baryon", "photon", "nu", "mass_nu", "total", "no_nu", "total_de", "Weyl",
                "v_cdm", "v_b", "v_b-v_cdm")


def read_camb_file(path: str) -> Tuple[np.ndarray, np.ndarray]:
    """(k [h/Mpc], total-matter T) of a CAMB transfer file, as
    transfer_CAMB_file.cc:36-100 reads it: lines holding '#' are skipped,
    every other line needs the 13 CAMB columns; the delta_matter column is
    'total'."""
    k, t = [], []
    with open(path) as f:
        for lineno, line in enumerate(f, 1):
            if "#" in line:
                continue
            cols = line.split()
            if not cols:
                continue
            try:
                if len(cols) < len(CAMB_COLUMNS):
                    raise ValueError
                vals = [float(x) for x in cols[:len(CAMB_COLUMNS)]]
            except ValueError:
                raise RuntimeError(f"error reading transfer function file '{path}' (line {lineno}): "
                                   "corrupt or not in the expected format") from None
            k.append(vals[0])
            t.append(vals[CAMB_COLUMNS.index("total")])
    if not k:
        raise RuntimeError(f"no transfer function rows in '{path}'")
    return np.array(k), np.array(t)


def with_table(p: L.CosmoParams, k, T) -> L.CosmoParams:
    """p with the tabulated transfer function (k in h/Mpc, increasing, T > 0).
    The arrays stay referenced by p for as long as p lives."""
    k = np.ascontiguousarray(k, dtype=np.float64)
    T = np.ascontiguousarray(T, dtype=np.float64)
    if k.shape != T.shape or k.ndim != 1:
        raise ValueError("k and T must be one-dimensional arrays of equal length")
    p._tk_arrays = (k, T)
    p.tk_k = k.ctypes.data_as(C.POINTER(C.c_double))
    p.tk_T = T.ctypes.data_as(C.POINTER(C.c_double))
    p.tk_n = k.size
    return p
