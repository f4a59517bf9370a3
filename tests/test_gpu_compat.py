"""include/mfic_compat.hpp on the device: the reference's LPT source block
(ic_generator.cc:294-659) written against Grid_FFT / RNG_ngenic / calculator
/ OrszagConvolver / op:: in C++ (tests/compat/compat_check.cpp) produces the
oracle's potentials, and Grid_FFT's element access and transforms agree with
numpy's unitary FFT on a non-cubic grid."""
import ctypes as C

import numpy as np
import pytest

from monofonic_amd import F32, F64
from monofonic_amd import lpt as P
from oracle import oracle as O

from _gpu_helpers import rel_l2
from test_cpu_compat import compat_lib

TOL = {F64: 1e-9, F32: 1e-5}


@pytest.fixture(scope="module")
def clib(ctx):
    lib = compat_lib()
    assert lib.compat_set_device(0) == 0
    yield lib
    assert lib.compat_release() == 0, lib.compat_last_error()


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [F64, F32])
@pytest.mark.parametrize("order", [2, 3])
def test_reference_source_block_matches_oracle(clib, prec, order):
    n, box, z, seed = 32, 64.0, 24.0, 4242
    count = 2 if order == 2 else 6
    out = np.zeros((count, n, n, n + 2))
    rc = clib.compat_potentials(n, box, seed, order, prec, C.byref(P.baseline_cosmology(z)), out.ctypes.data)
    assert rc == 0, clib.compat_last_error()
    phi, phi2, phi3, A3 = O.potentials(O.ic_cfg(n, box, order, seed), O.baseline_cosmology(z),
                                       O.ngenic_fill(seed, n))
    want = [phi, phi2] + ([phi3] + list(A3) if order == 3 else [])
    for i, w in enumerate(want):
        err = rel_l2(O.kview(out[i]), O.kview(w))
        print(f"compat potentials order {order} prec {prec} field {i}: rel-L2 {err:.3e}", flush=True)
        assert err < TOL[prec], (i, err)


@pytest.mark.gpu
def test_grid_element_access_and_transforms(clib):
    n0, n1, n2 = 24, 16, 36
    x = np.random.default_rng(7).standard_normal((n0, n1, n2))
    fwd = np.zeros((n0, n1, n2 // 2 + 1), dtype=np.complex128)
    back = np.zeros_like(x)
    rc = clib.compat_grid_roundtrip(n0, n1, n2, x.ctypes.data, fwd.ctypes.data, back.ctypes.data)
    assert rc == 0, clib.compat_last_error()
    assert rel_l2(fwd, np.fft.rfftn(x, norm="ortho")) < 1e-13
    assert rel_l2(back, x) < 1e-13
