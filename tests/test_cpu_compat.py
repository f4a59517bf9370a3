"""include/mfic_compat.hpp (the reference's C++ class interfaces over the C
ABI) without a GPU: its test library builds with g++ against libmfic.so and
loads, the cosmology calculator and E&H transfer wrappers equal the oracle,
and device work fails with an error (no CPU fallback)."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from monofonic_amd import _lib as L
from monofonic_amd import lpt as P
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SO = os.path.join(ROOT, "tests", "compat", "libcompat_check.so")
EXPORTS = ("compat_last_error", "compat_set_device", "compat_release", "compat_calculator",
           "compat_potentials", "compat_grid_roundtrip")


def compat_lib():
    if not os.path.exists(SO):
        L.load()
        subprocess.check_call(["make", "-s", "-C", os.path.dirname(SO)])
    lib = C.CDLL(SO)
    lib.compat_last_error.restype = C.c_char_p
    lib.compat_calculator.argtypes = [C.c_void_p, C.c_double, C.c_void_p]
    lib.compat_potentials.argtypes = [C.c_size_t, C.c_double, C.c_long, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
    lib.compat_grid_roundtrip.argtypes = [C.c_size_t] * 3 + [C.c_void_p] * 3
    return lib


def test_compat_library_exports():
    lib = compat_lib()
    assert all(hasattr(lib, n) for n in EXPORTS)


@pytest.mark.parametrize("z", [24.0, 49.0])
def test_calculator_equals_oracle(z):
    lib = compat_lib()
    p = P.baseline_cosmology(z)
    c = O.baseline_cosmology(z)
    coef = O.lpt_coefs(O.ic_cfg(64, 100.0, 1, 1), c)
    for k in (1e-3, 0.05, 0.7, 5.0):
        out = np.zeros(5)
        assert lib.compat_calculator(C.byref(p), k, out.ctypes.data) == 0, lib.compat_last_error()
        a = 1.0 / (1.0 + z)
        assert out[0] == coef.Dplus0
        assert out[0] == O.lib().orc_cosmo_D(c, a)
        assert out[2] == O.lib().orc_cosmo_f(c, a)
        assert np.isclose(out[1], coef.vfac, rtol=1e-15, atol=0)
        T = O.lib().orc_tk_eh(c, k)
        assert np.isclose(out[3], T, rtol=1e-14, atol=0)
        assert np.isclose(out[4], k ** (0.5 * c.n_s) * T * c.sqrtpnorm, rtol=1e-14, atol=0)


def test_calculator_refuses_tabulated_transfer():
    lib = compat_lib()
    p = P.baseline_cosmology(24.0)
    kk = np.array([1e-4, 1e-2, 1.0, 1e2])
    tt = np.ones(4)
    p.tk_k = kk.ctypes.data_as(C.POINTER(C.c_double))
    p.tk_T = tt.ctypes.data_as(C.POINTER(C.c_double))
    p.tk_n = 4
    assert lib.compat_calculator(C.byref(p), 0.1, np.zeros(5).ctypes.data) == -1
    assert b"tabulated" in lib.compat_last_error()


def test_device_work_fails_without_gpu():
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    lib = compat_lib()
    out = np.zeros(2 * 8 * 8 * 10)
    rc = lib.compat_potentials(8, 10.0, 1, 2, L.F64, C.byref(P.baseline_cosmology(24.0)), out.ctypes.data)
    assert rc == -1 and lib.compat_last_error().startswith(b"mfic: ")
