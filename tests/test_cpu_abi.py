"""The C-ABI library on a machine without a GPU: it loads, exports every
entry point that include/mfic.h declares, refuses device work with an error
code (never a crash or a CPU fallback), and its host-only entry points (seed
table, cosmology scalars, LPT coefficients) equal the oracle bit for bit."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from monofonic_amd import _lib as L
from monofonic_amd import lpt as P
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "mfic.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mfic_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_entry_point():
    lib = L.load()
    names = declared()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding types every declared entry point
    assert not [n for n in names if n not in L.SIGNATURES]
    assert lib.mfic_version().startswith(b"mfic")


def test_device_calls_fail_cleanly_without_a_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    h = C.c_void_p()
    rc = L.load().mfic_ctx_create(0, C.byref(h))
    assert rc == -3 and b"device" in L.load().mfic_last_error()
    with pytest.raises(L.MficError):
        L.call("mfic_ctx_create", 0, C.byref(h))
    assert L.load().mfic_grid_destroy(None) == 0 or L.load().mfic_grid_destroy(None) < 0


@pytest.mark.parametrize("n", [8, 64, 256])
def test_seedtable_equals_oracle(n):
    assert np.array_equal(P.seedtable(12345, n), O.seedtable(12345, n))


def test_cosmology_scalars_equal_oracle():
    for z in (24.0, 49.0):
        s = P.cosmo_scalars(P.baseline_cosmology(z))
        c = O.baseline_cosmology(z)
        for name in ("h", "Omega_nu_massive", "Omega_gamma", "Omega_r", "Omega_c", "Omega_DE", "Dnow",
                     "Dplus_start", "f_start", "vfac_start", "sigma8_measured", "pnorm", "sqrtpnorm"):
            assert getattr(s, name) == getattr(c, name), name
    p = P.baseline_cosmology(24.0)
    c = O.baseline_cosmology(24.0)
    a = np.array([0.01, 0.04, 0.3, 1.0])
    k = np.array([1e-3, 0.05, 0.7, 5.0])
    assert np.array_equal(P.cosmo_eval(p, "D", a), [O.lib().orc_cosmo_D(c, x) for x in a])
    assert np.array_equal(P.cosmo_eval(p, "f", a), [O.lib().orc_cosmo_f(c, x) for x in a])
    # T(k) goes through a different compiler's libm-call sequence: rounding-level agreement
    assert np.allclose(P.cosmo_eval(p, "T", k), [O.lib().orc_tk_eh(c, x) for x in k], rtol=1e-14, atol=0)


def test_lpt_coefficients_equal_oracle():
    p = P.baseline_cosmology(24.0)
    c = O.baseline_cosmology(24.0)
    for order in (1, 2, 3):
        got = P.lpt_coefficients(p, 256, 500.0, order)
        want = O.lpt_coefs(O.ic_cfg(256, 500.0, order, 1), c)
        for name in ("Dplus0", "vfac", "g1", "g2", "g3", "g3c", "vfac1", "vfac2", "vfac3", "volfac"):
            assert getattr(got, name) == getattr(want, name), (order, name)


def test_argument_checks():
    lib = L.load()
    assert lib.mfic_fft_size_supported(512) == 1
    assert lib.mfic_fft_size_supported(768) == 1
    assert lib.mfic_fft_size_supported(7) == 0
    assert lib.mfic_ngenic_seedtable(1, 8, None) == -1
    assert lib.mfic_cosmo_compute(None, None) == -1


def test_context_closes_children_before_itself(monkeypatch):
    """Context.close() closes every live handle owner it adopted before it
    calls mfic_ctx_destroy (which refuses while handles are alive)"""
    from monofonic_amd.grid import Context
    calls = []
    monkeypatch.setattr(L, "call", lambda name, *a: calls.append(name))

    class Child:
        def __init__(self, tag):
            self.tag = tag

        def close(self):
            calls.append(self.tag)

    ctx = Context.__new__(Context)
    ctx.h = C.c_void_p(0x1000)
    import weakref
    ctx._children = weakref.WeakSet()
    kids = [Child(f"child{i}") for i in range(3)]
    for k in kids:
        ctx.adopt(k)
    ctx.close()
    assert calls[-1] == "mfic_ctx_destroy"
    assert sorted(calls[:-1]) == ["child0", "child1", "child2"]
    assert not ctx.h


def test_radix_plans():
    """every compiled radix plan multiplies to its length, each radix divides
    the elements per thread, and a transform fits one wavefront group; the
    padded-pass lengths carry the alternative plan"""
    from monofonic_amd import batch
    sizes = [4, 6, 8, 9, 12, 16, 18, 24, 32, 36, 48, 64, 72, 96, 128, 144, 192, 256, 288, 384, 512, 576, 768]
    for length in sizes:
        count = batch.plan_count(length)
        assert count == (2 if length in (512, 768) else 1), length
        for v in range(count):
            e, radices = batch.plan_info(length, v)
            assert int(np.prod(radices)) == length and all(e % r == 0 for r in radices), (length, v)
            assert length % e == 0 and length // e <= 64
    assert batch.plan_info(768, 1) == (48, (16, 12, 4))
    assert batch.plan_info(512, 1) == (32, (16, 8, 4))
    assert batch.plan_count(7) == 0 and batch.plan_count(1024) == 0
    with pytest.raises(L.MficError):
        batch.plan_info(384, 1)
    # a context is needed to select a plan: without one the call fails cleanly
    assert L.load().mfic_fft_set_plan(None, 768, L.PASS_PAD_C2R_B, L.F32, 1) == -1
    v = C.c_int()
    assert L.load().mfic_fft_get_plan(None, 768, L.PASS_PAD_C2R_B, L.F32, C.byref(v)) == -1
