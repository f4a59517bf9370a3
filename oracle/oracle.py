"""ctypes binding of the CPU FP64 oracle (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this module, and only as the checker / CPU baseline.  The
product (``monofonic_amd``) never imports it.

Arrays follow the reference layout: a real field of logical size n^3 is a
float64 array of shape (n, n, n + 2); its k-space view is complex128 of shape
(n, n, n // 2 + 1) over the same memory (grid_fft.cc:45-54).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "build", "liboracle.so")
_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc)."""
    if force or not os.path.exists(_SO):
        subprocess.check_call(["make", "-s", "-C", _HERE] + (["-B"] if force else []))
    return _SO


class Cosmo(C.Structure):
    _fields_ = [
        ("Omega_m", C.c_double), ("Omega_b", C.c_double), ("H0", C.c_double),
        ("n_s", C.c_double), ("sigma_8", C.c_double), ("A_s", C.c_double), ("k_p", C.c_double),
        ("Tcmb", C.c_double), ("N_ur", C.c_double), ("m_nu", C.c_double * 3),
        ("deg_nu", C.c_double * 3), ("O_nu_norm", C.c_double),
        ("w_0", C.c_double), ("w_a", C.c_double), ("Omega_k", C.c_double),
        ("alpha_s", C.c_double), ("beta_s", C.c_double),
        ("zero_radiation", C.c_int),
        ("zstart", C.c_double), ("ztarget", C.c_double),
        ("h", C.c_double), ("Omega_nu_massive", C.c_double), ("Omega_gamma", C.c_double),
        ("Omega_r", C.c_double), ("Omega_c", C.c_double), ("Omega_DE", C.c_double),
        ("Dnow", C.c_double), ("Dplus_start", C.c_double), ("Dplus_target", C.c_double),
        ("vfac_start", C.c_double), ("f_start", C.c_double),
        ("sigma8_measured", C.c_double), ("pnorm", C.c_double), ("sqrtpnorm", C.c_double),
        ("eh", C.c_double * 24),
        ("ntab", C.c_int),
        ("tab_loga", C.c_void_p), ("tab_logD", C.c_void_p), ("tab_logf", C.c_void_p),
        ("spl_D", C.c_void_p), ("spl_f", C.c_void_p),
        ("tk_n", C.c_int), ("tk_k", C.c_void_p), ("tk_T", C.c_void_p),
        ("tk_lx", C.c_void_p), ("tk_ly", C.c_void_p), ("tk_m", C.c_void_p),
    ]


class IcCfg(C.Structure):
    _fields_ = [
        ("n", C.c_int), ("L", C.c_double), ("lpt_order", C.c_int), ("lattice", C.c_int),
        ("plt", C.c_int), ("do_fixing", C.c_int), ("do_inversion", C.c_int),
        ("seed", C.c_long), ("lunit", C.c_double), ("vunit", C.c_double),
    ]


class LptCoef(C.Structure):
    _fields_ = [(k, C.c_double) for k in
                ("Dplus0", "vfac", "g1", "g2", "g3", "g3c", "vfac1", "vfac2", "vfac3", "volfac")]


class Operand(C.Structure):
    _fields_ = [("f", C.c_void_p), ("kind", C.c_int), ("d", C.c_int * 4)]


class Plt(C.Structure):
    _fields_ = [("ng", C.c_int), ("Dr", C.POINTER(C.c_double)), ("Dtheta", C.POINTER(C.c_double)),
                ("Dphi", C.POINTER(C.c_double)), ("vfac", C.POINTER(C.c_double))]


class Ranlxd(C.Structure):
    _fields_ = [("x", C.c_uint64 * 12), ("carry", C.c_int), ("ir", C.c_int), ("jr", C.c_int),
                ("ir_old", C.c_int), ("pr", C.c_int)]


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(_SO)
        L.orc_fft3d_r2c.argtypes = [_dp, C.c_int, C.c_int, C.c_int]
        L.orc_fft3d_c2r.argtypes = [_dp, C.c_int, C.c_int, C.c_int]
        L.orc_fft1d.argtypes = [_dp, C.c_int, C.c_int]
        L.orc_fft3d_c2c.argtypes = [_dp, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_ppt_fields.argtypes = [C.POINTER(IcCfg), C.POINTER(LptCoef), _dp, C.c_void_p, C.c_double, _dp, _dp,
                                     _dp, _dp, C.POINTER(C.c_double)]
        L.orc_ppt_fields.restype = C.c_int
        L.orc_cr_log.restype = C.c_double
        L.orc_cr_log.argtypes = [C.c_double]
        L.orc_cr_sincos.argtypes = [C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_ranlxd_set.argtypes = [C.POINTER(Ranlxd), C.c_ulong, C.c_int]
        L.orc_ranlxd_uniform.restype = C.c_double
        L.orc_ranlxd_uniform.argtypes = [C.POINTER(Ranlxd)]
        L.orc_ranlxd_get.restype = C.c_ulong
        L.orc_ranlxd_get.argtypes = [C.POINTER(Ranlxd)]
        L.orc_ngenic_seedtable.argtypes = [C.c_long, C.c_size_t, _u32p]
        L.orc_ngenic_fill.argtypes = [_u32p, C.c_size_t, _dp]
        L.orc_ngenic_libm_compare.argtypes = [_u32p, C.c_size_t, C.POINTER(C.c_uint64 * 5)]
        L.orc_cosmo_defaults.argtypes = [C.POINTER(Cosmo)]
        L.orc_cosmo_init.argtypes = [C.POINTER(Cosmo)]
        L.orc_cosmo_init.restype = C.c_int
        L.orc_cosmo_free.argtypes = [C.POINTER(Cosmo)]
        for fn in ("orc_cosmo_H", "orc_cosmo_D", "orc_cosmo_f", "orc_cosmo_vfact", "orc_tk_eh", "orc_tk",
                   "orc_amplitude"):
            getattr(L, fn).restype = C.c_double
            getattr(L, fn).argtypes = [C.POINTER(Cosmo), C.c_double]
        L.orc_integrate_sigma8.restype = C.c_double
        L.orc_integrate_sigma8.argtypes = [C.POINTER(Cosmo), C.c_double, C.c_double, C.c_double]
        L.orc_cosmo_growth_table.restype = C.c_int
        L.orc_cosmo_growth_table.argtypes = [C.POINTER(Cosmo), _dp, _dp, _dp, C.c_int]
        L.orc_convolve2.argtypes = [C.c_int, C.c_double, C.POINTER(Operand), C.POINTER(Operand),
                                    _dp, C.c_int, C.c_double]
        L.orc_convolve3.argtypes = [C.c_int, C.c_double, C.POINTER(Operand), C.POINTER(Operand),
                                    C.POINTER(Operand), _dp, C.c_int, C.c_double]
        L.orc_lpt_coefs.argtypes = [C.POINTER(IcCfg), C.POINTER(Cosmo), C.POINTER(LptCoef)]
        L.orc_lpt_potentials.argtypes = [C.POINTER(IcCfg), C.POINTER(Cosmo), _dp] + [C.c_void_p] * 6
        L.orc_lpt_component_k.argtypes = [C.POINTER(IcCfg), C.POINTER(LptCoef), C.c_void_p] + \
            [C.c_void_p] * 6 + [C.c_int, _dp]
        L.orc_shift_field.argtypes = [C.c_int, C.c_double, C.POINTER(C.c_double * 3), _dp]
        L.orc_inverse_laplacian.argtypes = [C.c_int, C.c_double, _dp]
        L.orc_zero_dc.argtypes = [_dp]
        L.orc_scale.argtypes = [C.c_int, _dp, C.c_double]
        L.orc_lattice_set_positions.argtypes = [C.c_int, C.c_double, C.c_int, C.c_int, C.c_double, _dp, _dp]
        L.orc_lattice_readout_index.argtypes = [C.c_int, C.c_double, C.c_int, C.c_int, C.c_int, C.c_double,
                                                C.c_int, _dp, _dp]
        L.orc_lattice_set_velocities.argtypes = [C.c_int, C.c_double, C.c_int, C.c_int, _dp, _dp]
        L.orc_ic_run.argtypes = [C.POINTER(IcCfg), C.POINTER(Cosmo), C.c_void_p, C.c_void_p,
                                 C.c_void_p, C.c_void_p]
        L.orc_ic_run.restype = C.c_int
        L.orc_plt_init.argtypes = [C.POINTER(Plt), C.c_int, C.c_int]
        L.orc_plt_init.restype = C.c_int
        L.orc_plt_free.argtypes = [C.POINTER(Plt)]
        L.orc_plt_dynmat.argtypes = [C.c_int, C.c_int, _dp]
        L.orc_plt_gradient.restype = C.c_double
        L.orc_plt_gradient.argtypes = [C.POINTER(Plt), C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int]
        L.orc_plt_vfac_corr.restype = C.c_double
        L.orc_plt_vfac_corr.argtypes = [C.POINTER(Plt), C.c_int, C.c_int, C.c_int, C.c_int]
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data


# ---------------------------------------------------------------- helpers --
def empty_field(n: int) -> np.ndarray:
    return np.zeros((n, n, n + 2), dtype=np.float64)


def kview(a: np.ndarray) -> np.ndarray:
    """complex k-space view (n, n, n//2+1) of an (n, n, n+2) float64 array"""
    return a.view(np.complex128)


def rview(a: np.ndarray) -> np.ndarray:
    """logical real-space view (n, n, n) of an (n, n, n+2) float64 array"""
    return a[:, :, : a.shape[1]]


def fft_r2c(a: np.ndarray) -> None:
    n0, n1, n2p = a.shape
    lib().orc_fft3d_r2c(a, n0, n1, n2p - 2)


def fft_c2r(a: np.ndarray) -> None:
    n0, n1, n2p = a.shape
    lib().orc_fft3d_c2r(a, n0, n1, n2p - 2)


def fft1d(z: np.ndarray, sign: int) -> np.ndarray:
    buf = np.ascontiguousarray(z, dtype=np.complex128).copy()
    lib().orc_fft1d(buf.view(np.float64), buf.size, sign)
    return buf


def fft_c2c(z: np.ndarray, sign: int) -> np.ndarray:
    """unitary 3D c2c transform of a complex (n0, n1, n2) array (copy)"""
    buf = np.ascontiguousarray(z, dtype=np.complex128).copy()
    n0, n1, n2 = buf.shape
    lib().orc_fft3d_c2c(buf.view(np.float64).reshape(-1), n0, n1, n2, sign)
    return buf


def ppt_fields(cfg, coef, phi_k: np.ndarray, phi2_k, velocity_unit: float):
    """semiclassical PT fields: (delta, vx, vy, vz) real (n, n, n+2) arrays and hbar"""
    n = cfg.n
    outs = [empty_field(n) for _ in range(4)]
    hb = C.c_double()
    rc = lib().orc_ppt_fields(C.byref(cfg), C.byref(coef), phi_k, _ptr(phi2_k), velocity_unit, *outs, C.byref(hb))
    if rc != 0:
        raise RuntimeError("oracle PPT failed")
    return outs, hb.value


def cr_log(x: float) -> float:
    return lib().orc_cr_log(x)


def cr_sincos(x: float):
    s, c = C.c_double(), C.c_double()
    lib().orc_cr_sincos(x, C.byref(s), C.byref(c))
    return s.value, c.value


class Ranlux:
    def __init__(self, seed: int, luxury: int = 202):
        self.st = Ranlxd()
        lib().orc_ranlxd_set(C.byref(self.st), seed, luxury)

    def uniform(self) -> float:
        return lib().orc_ranlxd_uniform(C.byref(self.st))

    def get(self) -> int:
        return lib().orc_ranlxd_get(C.byref(self.st))


def seedtable(seed: int, n: int) -> np.ndarray:
    t = np.zeros(n * n, dtype=np.uint32)
    lib().orc_ngenic_seedtable(seed, n, t)
    return t


def ngenic_fill(seed: int, n: int) -> np.ndarray:
    t = seedtable(seed, n)
    out = empty_field(n)
    lib().orc_ngenic_fill(t, n, out)
    return out


def ngenic_libm_compare(seed: int, n: int) -> dict:
    """counts of draws whose host-libm Gaussian transform (the reference's
    std::log/cos/sin) differs from the correctly rounded one"""
    out = (C.c_uint64 * 5)()
    lib().orc_ngenic_libm_compare(seedtable(seed, n), n, C.byref(out))
    return dict(zip(("draws", "log", "cos", "sin", "value"), [int(x) for x in out]))


def cosmology(transfer=None, **kw) -> Cosmo:
    """Cosmology with the reference defaults (ParameterSet=none) and overrides.
    Giving sigma_8 selects sigma_8 normalisation (A_s = -1), as the reference.
    transfer=(k, T): a tabulated transfer function (k in h/Mpc), kept alive
    on the returned object."""
    c = Cosmo()
    lib().orc_cosmo_defaults(C.byref(c))
    if transfer is not None:
        k = np.ascontiguousarray(transfer[0], dtype=np.float64)
        T = np.ascontiguousarray(transfer[1], dtype=np.float64)
        c._tk_arrays = (k, T)
        c.tk_n = len(k)
        c.tk_k = k.ctypes.data
        c.tk_T = T.ctypes.data
    for k, v in kw.items():
        if k in ("m_nu", "deg_nu"):
            for i, x in enumerate(v):
                getattr(c, k)[i] = x
        else:
            setattr(c, k, v)
    if "sigma_8" in kw and "A_s" not in kw:
        c.A_s = -1.0
    rc = lib().orc_cosmo_init(C.byref(c))
    if rc != 0:
        raise RuntimeError("oracle cosmology initialisation failed")
    return c


def baseline_cosmology(zstart: float) -> Cosmo:
    """The cosmology shared by every benchmark configuration."""
    return cosmology(Omega_m=0.3, Omega_b=0.045, H0=70.0, sigma_8=0.81, n_s=0.961, zstart=zstart)


def operand(field: np.ndarray, kind: int, d=(0, 0, 0, 0)) -> Operand:
    o = Operand()
    o.f = field.ctypes.data
    o.kind = kind
    for i in range(4):
        o.d[i] = d[i] if i < len(d) else 0
    return o


def convolve2(n, L, a: Operand, b: Operand, out: np.ndarray, op: int, coef: float = 0.0):
    lib().orc_convolve2(n, L, C.byref(a), C.byref(b), out, op, coef)


def convolve3(n, L, a: Operand, b: Operand, c: Operand, out: np.ndarray, op: int, coef: float = 0.0):
    lib().orc_convolve3(n, L, C.byref(a), C.byref(b), C.byref(c), out, op, coef)


def ic_cfg(n, L, lpt_order, seed, lattice=0, plt=0, lunit=1.0, vunit=1.0, fixing=0, inversion=0):
    cfg = IcCfg()
    cfg.n, cfg.L, cfg.lpt_order, cfg.lattice, cfg.plt = n, L, lpt_order, lattice, plt
    cfg.do_fixing, cfg.do_inversion, cfg.seed, cfg.lunit, cfg.vunit = fixing, inversion, seed, lunit, vunit
    return cfg


def lpt_coefs(cfg: IcCfg, cosmo: Cosmo) -> LptCoef:
    k = LptCoef()
    lib().orc_lpt_coefs(C.byref(cfg), C.byref(cosmo), C.byref(k))
    return k


def potentials(cfg: IcCfg, cosmo: Cosmo, wnoise_k: np.ndarray):
    n = cfg.n
    phi = empty_field(n)
    phi2 = empty_field(n) if cfg.lpt_order > 1 else None
    rest = [empty_field(n) for _ in range(4)] if cfg.lpt_order > 2 else [None] * 4
    lib().orc_lpt_potentials(C.byref(cfg), C.byref(cosmo), wnoise_k, _ptr(phi), _ptr(phi2),
                             *[_ptr(x) for x in rest])
    return phi, phi2, rest[0], rest[1:]


def component_k(cfg, coef, pots, comp, plt=None):
    phi, phi2, phi3, A3 = pots
    out = empty_field(cfg.n)
    lib().orc_lpt_component_k(C.byref(cfg), C.byref(coef), None if plt is None else C.byref(plt),
                              _ptr(phi), _ptr(phi2), _ptr(phi3), *[_ptr(a) for a in A3], comp, out)
    return out


def ic_run(cfg: IcCfg, cosmo: Cosmo, particles: bool = False):
    """Whole pipeline. Returns (fields6 (6,n,n,n+2), pos, vel, stage_seconds)."""
    n = cfg.n
    fields = np.zeros((6, n, n, n + 2))
    pos = vel = None
    if particles:
        npart = (1 << cfg.lattice) * n ** 3
        pos = np.zeros(3 * npart)
        vel = np.zeros(3 * npart)
    st = np.zeros(8)
    rc = lib().orc_ic_run(C.byref(cfg), C.byref(cosmo), fields.ctypes.data, _ptr(pos), _ptr(vel),
                          st.ctypes.data)
    if rc != 0:
        raise RuntimeError("oracle ic_run failed")
    return fields, pos, vel, st


def shift_field(a: np.ndarray, L: float, s) -> None:
    """shift_field in place on an (n, n, n+2) real-space array"""
    lib().orc_shift_field(a.shape[0], L, C.byref((C.c_double * 3)(*s)), a)


def inverse_laplacian(a: np.ndarray, L: float) -> None:
    lib().orc_inverse_laplacian(a.shape[0], L, a)


def zero_dc(a: np.ndarray) -> None:
    lib().orc_zero_dc(a)


def scale(a: np.ndarray, s: float) -> None:
    lib().orc_scale(a.shape[0], a, s)


def lattice_readout(field: np.ndarray, L: float, lattice: int, idim: int, is_pos: bool, lunit: float = 1.0):
    """set_positions / set_velocities of one component (field is shifted in
    place for bcc); returns the (npart, 3) array with column idim filled"""
    n = field.shape[0]
    out = np.zeros(3 * (1 << lattice) * n ** 3)
    if is_pos:
        lib().orc_lattice_set_positions(n, L, lattice, idim, lunit, field, out)
    else:
        lib().orc_lattice_set_velocities(n, L, lattice, idim, field, out)
    return out.reshape(-1, 3)


def lattice_readout_index(field: np.ndarray, L: float, lattice: int, lattice_index: int, idim: int, is_pos: bool,
                          lunit: float = 1.0):
    """readout of one component (real-space field, modified in place as the
    reference modifies it) on lattice index 0 or 1; (npart, 3) array with
    column idim set"""
    n = field.shape[0]
    out = np.zeros(((1 << lattice) * n ** 3, 3))
    lib().orc_lattice_readout_index(n, L, lattice, lattice_index, idim, lunit, int(is_pos), field, out)
    return out


def plt_tables(lattice: int, ng: int = 64) -> Plt:
    p = Plt()
    if lib().orc_plt_init(C.byref(p), lattice, ng) != 0:
        raise RuntimeError("PLT init failed")
    return p


def plt_dynmat(lattice: int, ng: int) -> np.ndarray:
    """(6, ng, ng, ng/2+1): D_xx D_xy D_xz D_yy D_yz D_zz in k-space"""
    out = np.zeros(6 * ng * ng * (ng // 2 + 1))
    lib().orc_plt_dynmat(lattice, ng, out)
    return out.reshape(6, ng, ng, ng // 2 + 1)


def plt_arrays(p: Plt):
    ng = p.ng
    shape = (ng, ng, ng // 2 + 1)
    cnt = ng * ng * (ng // 2 + 1)
    return {name: np.ctypeslib.as_array(getattr(p, name), shape=(cnt,)).reshape(shape).copy()
            for name in ("Dr", "Dtheta", "Dphi", "vfac")}


# operand kinds / output ops (mirror oracle.h)
HESS, SUM_HESS, DIFF_HESS, GRAD, IDENT = 0, 1, 2, 3, 4
ASSIGN, ADD, SUB, MADD = 0, 1, 2, 3
