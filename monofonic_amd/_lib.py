"""ctypes binding of libmfic.so (the HIP/gfx950 C ABI declared in include/mfic.h).

The library is built in-tree by ``__graft_entry__.build()`` (or
``make -C monofonic_amd/csrc``).  There is no CPU fallback: if the shared
library is missing or fails to load, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmfic.so")

F32, F64 = 0, 1
KSPACE, RSPACE = 0, 1
OP_ASSIGN, OP_ADD, OP_SUB, OP_MADD = 0, 1, 2, 3
HESS, SUM_HESS, DIFF_HESS, GRAD, IDENT = 0, 1, 2, 3, 4
SC, BCC, FCC = 0, 1, 2
# strided pass kinds with a selectable radix plan (mfic_fft_pass)
PASS_PENCILS, PASS_AX0_SHIFT = 0, 2
PASS_PAD_C2R_A, PASS_PAD_C2R_B, PASS_PAD_R2C_A, PASS_PAD_R2C_B, PASS_TRUNC_AX0, PASS_TRUNC_AX1 = 3, 4, 5, 6, 7, 8
# fused source row passes: form 0 = two rows per full-length complex FFT, 1 = one half-length FFT per row
PASS_FUSED_ROWS = 9
PASSES = (PASS_PENCILS, PASS_AX0_SHIFT, PASS_PAD_C2R_A, PASS_PAD_C2R_B, PASS_PAD_R2C_A, PASS_PAD_R2C_B,
          PASS_TRUNC_AX0, PASS_TRUNC_AX1, PASS_FUSED_ROWS)

ERRORS = {
    -1: "MFIC_ERR_ARG", -2: "MFIC_ERR_SIZE", -3: "MFIC_ERR_HIP", -4: "MFIC_ERR_ALLOC",
    -5: "MFIC_ERR_STATE", -6: "MFIC_ERR_CONFIG", -7: "MFIC_ERR_UNSUPPORTED", -8: "MFIC_ERR_IO",
}


class MficError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class CosmoParams(C.Structure):
    _fields_ = [
        ("Omega_m", C.c_double), ("Omega_b", C.c_double), ("H0", C.c_double), ("n_s", C.c_double),
        ("sigma_8", C.c_double), ("A_s", C.c_double), ("k_p", C.c_double),
        ("Tcmb", C.c_double), ("N_ur", C.c_double), ("m_nu", C.c_double * 3), ("deg_nu", C.c_double * 3),
        ("O_nu_norm", C.c_double), ("w_0", C.c_double), ("w_a", C.c_double), ("Omega_k", C.c_double),
        ("alpha_s", C.c_double), ("beta_s", C.c_double), ("zero_radiation", C.c_int),
        ("zstart", C.c_double), ("ztarget", C.c_double),
        ("tk_k", C.POINTER(C.c_double)), ("tk_T", C.POINTER(C.c_double)), ("tk_n", C.c_size_t),
    ]


class CosmoScalars(C.Structure):
    _fields_ = [(k, C.c_double) for k in (
        "h", "Omega_nu_massive", "Omega_gamma", "Omega_r", "Omega_c", "Omega_DE", "Dnow", "Dplus_start",
        "Dplus_target", "f_start", "vfac_start", "sigma8_measured", "pnorm", "sqrtpnorm")]


class LptCoef(C.Structure):
    _fields_ = [(k, C.c_double) for k in
                ("Dplus0", "vfac", "g1", "g2", "g3", "g3c", "vfac1", "vfac2", "vfac3", "volfac")]


class Operand(C.Structure):
    _fields_ = [("f", C.c_void_p), ("kind", C.c_int), ("d", C.c_int * 4)]


class IcConfig(C.Structure):
    _fields_ = [("n", C.c_size_t), ("L", C.c_double), ("lpt_order", C.c_int), ("lattice", C.c_int),
                ("plt", C.c_int), ("prec", C.c_int), ("do_fixing", C.c_int), ("do_inversion", C.c_int),
                ("seed", C.c_long), ("particles", C.c_int), ("lunit", C.c_double), ("vunit", C.c_double),
                ("baryons", C.c_int)]


_P = C.c_void_p
_sz3 = C.c_size_t * 3
_d3 = C.c_double * 3

# name -> (restype, argtypes); every int-returning entry point is status-checked
class GadgetOpts(C.Structure):
    _fields_ = [("double_precision", C.c_int), ("long_ids", C.c_int), ("gadget2_compat", C.c_int)]


class SwiftOpts(C.Structure):
    _fields_ = [("long_ids", C.c_int), ("YHe", C.c_double), ("gamma", C.c_double)]


SIGNATURES = {
    "mfic_last_error": (C.c_char_p, []),
    "mfic_version": (C.c_char_p, []),
    "mfic_ctx_create": (C.c_int, [C.c_int, C.POINTER(_P)]),
    "mfic_ctx_destroy": (C.c_int, [_P]),
    "mfic_ctx_sync": (C.c_int, [_P]),
    "mfic_ctx_stream": (C.c_int, [_P, C.POINTER(_P)]),
    "mfic_prof_enable": (C.c_int, [_P, C.c_int]),
    "mfic_prof_report": (C.c_int, [_P, C.c_char_p, C.c_size_t]),
    "mfic_grid_create": (C.c_int, [_P, C.POINTER(C.c_size_t), C.POINTER(C.c_double), C.c_int, C.c_int, C.POINTER(_P)]),
    "mfic_grid_destroy": (C.c_int, [_P]),
    "mfic_grid_set_space": (C.c_int, [_P, C.c_int]),
    "mfic_grid_get_space": (C.c_int, [_P, C.POINTER(C.c_int)]),
    "mfic_grid_device_ptr": (C.c_int, [_P, C.POINTER(_P), C.POINTER(C.c_size_t)]),
    "mfic_grid_upload": (C.c_int, [_P, _P, C.c_int]),
    "mfic_grid_download": (C.c_int, [_P, _P, C.c_int]),
    "mfic_grid_copy": (C.c_int, [_P, _P]),
    "mfic_grid_zero": (C.c_int, [_P]),
    "mfic_grid_scale": (C.c_int, [_P, C.c_double]),
    "mfic_grid_zero_dc": (C.c_int, [_P]),
    "mfic_grid_inverse_laplacian": (C.c_int, [_P]),
    "mfic_grid_shift": (C.c_int, [_P, C.POINTER(C.c_double), C.c_int]),
    "mfic_fft_forward": (C.c_int, [_P]),
    "mfic_fft_backward": (C.c_int, [_P]),
    "mfic_fft_size_supported": (C.c_int, [C.c_size_t]),
    "mfic_fft_set_tile": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    "mfic_fft_get_tile": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "mfic_fft_plan_count": (C.c_int, [C.c_int]),
    "mfic_fft_plan_info": (C.c_int, [C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "mfic_fft_set_plan": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_int]),
    "mfic_fft_get_plan": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int)]),
    "mfic_ngenic_seedtable": (C.c_int, [C.c_long, C.c_size_t, _P]),
    "mfic_ngenic_fill": (C.c_int, [_P, C.c_long]),
    "mfic_cosmo_defaults": (None, [C.POINTER(CosmoParams)]),
    "mfic_cosmo_compute": (C.c_int, [C.POINTER(CosmoParams), C.POINTER(CosmoScalars)]),
    "mfic_cosmo_eval": (C.c_int, [C.POINTER(CosmoParams), C.c_char_p, _P, _P, C.c_size_t]),
    "mfic_phi1_from_noise": (C.c_int, [_P, _P, C.POINTER(CosmoParams), C.c_int, C.c_int]),
    "mfic_phi1_from_noise_noalloc": (C.c_int, [_P, _P, C.POINTER(CosmoParams), C.c_int, C.c_int]),
    "mfic_conv_create": (C.c_int, [_P, C.POINTER(C.c_size_t), C.POINTER(C.c_double), C.c_int, C.POINTER(_P)]),
    "mfic_conv_destroy": (C.c_int, [_P]),
    "mfic_conv_convolve2": (C.c_int, [_P, C.POINTER(Operand), C.POINTER(Operand), _P, C.c_int, C.c_double]),
    "mfic_conv_convolve3": (C.c_int, [_P, C.POINTER(Operand), C.POINTER(Operand), C.POINTER(Operand), _P,
                                      C.c_int, C.c_double]),
    "mfic_lpt_sources": (C.c_int, [_P, C.c_int] + [_P] * 6),
    "mfic_lpt_coefficients": (C.c_int, [C.POINTER(CosmoParams), C.c_size_t, C.c_double, C.c_int,
                                        C.POINTER(LptCoef)]),
    "mfic_plt_create": (C.c_int, [_P, C.c_int, C.c_int, C.POINTER(_P)]),
    "mfic_plt_destroy": (C.c_int, [_P]),
    "mfic_plt_download": (C.c_int, [_P, _P, _P, _P, _P]),
    "mfic_lpt_component": (C.c_int, [_P, C.c_int, C.c_int, C.POINTER(LptCoef), _P, C.c_double] + [_P] * 7),
    "mfic_lattice_readout": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_double, _P]),
    "mfic_lattice_readout_index": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, _P]),
    "mfic_run_create": (C.c_int, [_P, C.POINTER(IcConfig), C.POINTER(CosmoParams), C.POINTER(_P)]),
    "mfic_fft_c2c": (C.c_int, [_P, _P, C.POINTER(C.c_size_t), C.c_int, C.c_int]),
    "mfic_ppt_fields": (C.c_int, [_P, C.c_double, _P, _P, _P, _P, C.POINTER(C.c_double)]),
    "mfic_gadget_units": (C.c_int, [C.c_double, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                    C.POINTER(C.c_double)]),
    "mfic_gadget_write": (C.c_int, [_P, C.c_char_p, C.POINTER(GadgetOpts)]),
    "mfic_gadget_write_host": (C.c_int, [C.c_char_p, C.POINTER(IcConfig), C.POINTER(CosmoParams),
                                         C.POINTER(GadgetOpts), _P, _P]),
    "mfic_swift_units": (C.c_int, [C.c_double, C.POINTER(CosmoParams), C.POINTER(C.c_double),
                                   C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "mfic_swift_write": (C.c_int, [_P, C.c_char_p, C.POINTER(SwiftOpts)]),
    "mfic_swift_write_host": (C.c_int, [C.c_char_p, C.POINTER(IcConfig), C.POINTER(CosmoParams),
                                        C.POINTER(SwiftOpts), _P, _P]),
    "mfic_run_destroy": (C.c_int, [_P]),
    "mfic_run_execute": (C.c_int, [_P, C.c_int]),
    "mfic_run_field": (C.c_int, [_P, C.c_int, C.POINTER(_P)]),
    "mfic_run_particles": (C.c_int, [_P, C.POINTER(_P), C.POINTER(_P), C.POINTER(C.c_size_t)]),
    "mfic_run_baryon_particles": (C.c_int, [_P, C.POINTER(_P), C.POINTER(_P), C.POINTER(C.c_size_t)]),
    "mfic_run_potential": (C.c_int, [_P, C.c_int, C.POINTER(_P)]),
    "mfic_graph_begin": (C.c_int, [_P]),
    "mfic_graph_end": (C.c_int, [_P, C.POINTER(_P)]),
    "mfic_graph_launch": (C.c_int, [_P, C.c_int]),
    "mfic_graph_destroy": (C.c_int, [_P]),
    "mfic_autotune": (C.c_int, [_P, C.c_size_t, C.c_int, C.c_int, C.POINTER(C.c_double)]),
    "mfic_autotune_config": (C.c_int, [_P, C.POINTER(IcConfig), C.POINTER(CosmoParams), C.POINTER(C.c_double)]),
    "mfic_lpt2_create": (C.c_int, [_P, C.c_size_t, C.c_double, C.c_int, C.POINTER(CosmoParams), C.c_double,
                                   C.POINTER(_P)]),
    "mfic_lpt2_destroy": (C.c_int, [_P]),
    "mfic_lpt2_forward": (C.c_int, [_P] * 5),
    "mfic_lpt2_adjoint": (C.c_int, [_P] * 6),
    "mfic_grid_wrap": (C.c_int, [_P, C.POINTER(C.c_size_t), C.POINTER(C.c_double), C.c_int, C.c_int, _P,
                                 C.POINTER(_P)]),
}

_lib = None


def load() -> C.CDLL:
    """Load libmfic.so; raises if it is missing (no fallback path exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build the HIP library first (python -c "
            "'import __graft_entry__ as g; g.build()' or make -C monofonic_amd/csrc)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def exported_symbols() -> list:
    lib = load()
    return [n for n in SIGNATURES if hasattr(lib, n)]


def check(rc: int) -> None:
    if rc != 0:
        msg = load().mfic_last_error()
        raise MficError(rc, msg.decode() if msg else "")


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args))
