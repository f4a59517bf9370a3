"""Host-side mirrors of the reference's convolver, LPT and run interfaces.

* ``OrszagConvolver``  -- convolution.hh:412-651 (convolve_Hessians & co.)
* ``lpt_sources``      -- ic_generator.cc:523-629 (reorganised schedule)
* ``ICRun``            -- the hot path of ic_generator::run
                          (ic_generator.cc:246-1060) for one box
* ``cosmo_params``     -- cosmology::parameters with ParameterSet=none
All compute goes through libmfic.so (HIP, gfx950); nothing here computes on
the CPU besides argument marshalling.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

from . import _lib as L
from .grid import Context, Grid


def cosmo_params(**kw) -> L.CosmoParams:
    """Reference defaults (cosmology_parameters.cc:110-133) plus overrides.
    Passing sigma_8 selects sigma_8 normalisation (A_s = -1), as the reference."""
    p = L.CosmoParams()
    L.load().mfic_cosmo_defaults(C.byref(p))
    for k, v in kw.items():
        if k in ("m_nu", "deg_nu"):
            for i, x in enumerate(v):
                getattr(p, k)[i] = x
        else:
            setattr(p, k, v)
    if "sigma_8" in kw and "A_s" not in kw:
        p.A_s = -1.0
    return p


def baseline_cosmology(zstart: float) -> L.CosmoParams:
    """Omega_m 0.3, Omega_b 0.045, h 0.7, sigma_8 0.81, n_s 0.961 (E&H)."""
    return cosmo_params(Omega_m=0.3, Omega_b=0.045, H0=70.0, sigma_8=0.81, n_s=0.961, zstart=zstart)


def cosmo_scalars(p: L.CosmoParams) -> L.CosmoScalars:
    s = L.CosmoScalars()
    L.call("mfic_cosmo_compute", C.byref(p), C.byref(s))
    return s


def cosmo_eval(p: L.CosmoParams, what: str, x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(x)
    L.call("mfic_cosmo_eval", C.byref(p), what.encode(), x.ctypes.data, y.ctypes.data, x.size)
    return y


def lpt_coefficients(p: L.CosmoParams, n: int, box: float, order: int) -> L.LptCoef:
    k = L.LptCoef()
    L.call("mfic_lpt_coefficients", C.byref(p), n, box, order, C.byref(k))
    return k


def seedtable(seed: int, n: int) -> np.ndarray:
    t = np.zeros(n * n, dtype=np.uint32)
    L.call("mfic_ngenic_seedtable", seed, n, t.ctypes.data)
    return t


def ngenic_fill(grid: Grid, seed: int) -> None:
    """RNG_ngenic::Fill_Grid on the device (random_ngenic.cc:84-154)."""
    L.call("mfic_ngenic_fill", grid.h, seed)


def phi1_from_noise(phi: Grid, wn: Grid, p: L.CosmoParams, fixing: bool = False, inversion: bool = False) -> None:
    L.call("mfic_phi1_from_noise", phi.h, wn.h, C.byref(p), int(fixing), int(inversion))


def _operand(g: Grid, kind: int, d: Sequence[int]) -> L.Operand:
    o = L.Operand()
    o.f = g.h
    o.kind = kind
    dd = list(d) + [0] * (4 - len(d))
    for i in range(4):
        o.d[i] = dd[i]
    return o


class OrszagConvolver:
    """OrszagConvolver<real_t>(N, L) with the BaseConvolver entry points."""

    def __init__(self, ctx: Context, n: Sequence[int], length: Sequence[float], prec: int = L.F32):
        self.h = C.c_void_p()
        L.call("mfic_conv_create", ctx.h, (C.c_size_t * 3)(*n), (C.c_double * 3)(*length), prec, C.byref(self.h))
        ctx.adopt(self)

    @staticmethod
    def _op(op):
        return op if isinstance(op, tuple) else (op, 0.0)

    def convolve2(self, a: L.Operand, b: L.Operand, out: Grid, op=L.OP_ASSIGN) -> None:
        o, c = self._op(op)
        L.call("mfic_conv_convolve2", self.h, C.byref(a), C.byref(b), out.h, o, c)

    def convolve3(self, a, b, c, out: Grid, op=L.OP_ASSIGN) -> None:
        o, cf = self._op(op)
        L.call("mfic_conv_convolve3", self.h, C.byref(a), C.byref(b), C.byref(c), out.h, o, cf)

    # convolution.hh:137-160, 171-209, 222-239, 258-274, 74-96
    def convolve_Hessians(self, inl: Grid, d2l, inr: Grid, d2r, *rest):
        if len(rest) == 1:
            self.convolve2(_operand(inl, L.HESS, d2l), _operand(inr, L.HESS, d2r), rest[0][0], rest[0][1])
        else:
            inm, d2m, op = rest[0], rest[1], rest[2]
            self.convolve3(_operand(inl, L.HESS, d2l), _operand(inr, L.HESS, d2r), _operand(inm, L.HESS, d2m),
                           op[0], op[1])

    def convolve_SumOfHessians(self, inl: Grid, d2l, inr: Grid, d2r1, d2r2, op):
        self.convolve2(_operand(inl, L.HESS, d2l), _operand(inr, L.SUM_HESS, list(d2r1) + list(d2r2)), op[0], op[1])

    def convolve_DifferenceOfHessians(self, inl: Grid, d2l, inr: Grid, d2r1, d2r2, op):
        self.convolve2(_operand(inl, L.HESS, d2l), _operand(inr, L.DIFF_HESS, list(d2r1) + list(d2r2)), op[0], op[1])

    def convolve_Gradients(self, inl: Grid, d1l, inr: Grid, d1r, op):
        self.convolve2(_operand(inl, L.GRAD, d1l), _operand(inr, L.GRAD, d1r), op[0], op[1])

    def close(self):
        if self.h:
            L.call("mfic_conv_destroy", self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# output operators (operators.hh:25-35): (grid, (op, coef)) pairs
def assign_to(g: Grid):
    return (g, (L.OP_ASSIGN, 0.0))


def add_to(g: Grid):
    return (g, (L.OP_ADD, 0.0))


def subtract_from(g: Grid):
    return (g, (L.OP_SUB, 0.0))


def multiply_add_to(g: Grid, x: float):
    return (g, (L.OP_MADD, float(x)))


def lpt_sources(ctx: Context, order: int, phi1: Grid, phi2: Optional[Grid] = None, phi3: Optional[Grid] = None,
                A3: Sequence[Optional[Grid]] = (None, None, None)) -> None:
    h = lambda g: None if g is None else g.h  # noqa: E731
    L.call("mfic_lpt_sources", ctx.h, order, phi1.h, h(phi2), h(phi3), h(A3[0]), h(A3[1]), h(A3[2]))


class PltTables:
    """particle::lattice_gradient tables (particle_plt.hh:520-549)."""

    def __init__(self, ctx: Context, lattice: int, ng: int = 64):
        self.h = C.c_void_p()
        self.ng = ng
        L.call("mfic_plt_create", ctx.h, lattice, ng, C.byref(self.h))
        ctx.adopt(self)

    def download(self):
        nk = self.ng * self.ng * (self.ng // 2 + 1)
        out = [np.empty(nk) for _ in range(4)]
        L.call("mfic_plt_download", self.h, *[o.ctypes.data for o in out])
        shape = (self.ng, self.ng, self.ng // 2 + 1)
        return {k: o.reshape(shape) for k, o in zip(("Dr", "Dtheta", "Dphi", "vfac"), out)}

    def close(self):
        if self.h:
            L.call("mfic_plt_destroy", self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def lpt_component(ctx: Context, order: int, comp: int, coef: L.LptCoef, pots, out: Grid, unit: float,
                  plt: Optional[PltTables] = None) -> None:
    h = lambda g: None if g is None else g.h  # noqa: E731
    phi1, phi2, phi3, A3 = pots
    L.call("mfic_lpt_component", ctx.h, order, comp, C.byref(coef), None if plt is None else plt.h, unit,
           phi1.h, h(phi2), h(phi3), h(A3[0]), h(A3[1]), h(A3[2]), out.h)


@dataclass
class ICConfig:
    """The [setup] / [random] / [output] keys the hot path reads
    (ic_generator.cc:100-137)."""
    n: int
    box: float
    zstart: float
    lpt_order: int = 3
    lattice: int = L.SC
    plt: bool = False
    prec: int = L.F32
    seed: int = 12345
    particles: bool = False
    fixing: bool = False
    inversion: bool = False
    lunit: float = 1.0
    vunit: float = 1.0
    baryons: bool = False

    def to_c(self) -> L.IcConfig:
        c = L.IcConfig()
        c.n, c.L, c.lpt_order, c.lattice, c.plt, c.prec = self.n, self.box, self.lpt_order, self.lattice, \
            int(self.plt), self.prec
        c.do_fixing, c.do_inversion, c.seed, c.particles = int(self.fixing), int(self.inversion), self.seed, \
            int(self.particles)
        c.lunit, c.vunit, c.baryons = self.lunit, self.vunit, int(self.baryons)
        return c


class ICRun:
    """One box of ic_generator::run on the device (mfic_run)."""

    def __init__(self, ctx: Context, cfg: ICConfig, cosmo: Optional[L.CosmoParams] = None):
        self.ctx, self.cfg = ctx, cfg
        self.cosmo = cosmo if cosmo is not None else baseline_cosmology(cfg.zstart)
        self.cosmo.zstart = cfg.zstart
        self._c = cfg.to_c()
        self.h = C.c_void_p()
        L.call("mfic_run_create", ctx.h, C.byref(self._c), C.byref(self.cosmo), C.byref(self.h))
        ctx.adopt(self)

    def execute(self, blocking: bool = True) -> None:
        L.call("mfic_run_execute", self.h, int(blocking))

    def field(self, comp: int) -> Grid:
        g = C.c_void_p()
        L.call("mfic_run_field", self.h, comp, C.byref(g))
        return Grid(self.ctx, (self.cfg.n,) * 3, (self.cfg.box,) * 3, prec=self.cfg.prec, handle=g)

    def potential(self, which: int) -> Grid:
        g = C.c_void_p()
        L.call("mfic_run_potential", self.h, which, C.byref(g))
        return Grid(self.ctx, (self.cfg.n,) * 3, (self.cfg.box,) * 3, prec=self.cfg.prec, handle=g)

    def particles_device(self, baryons: bool = False):
        """device pointers of the dark matter (or, with baryons=True, the
        baryon species') FP64 [p][3] positions and velocities, and npart"""
        pos, vel, n = C.c_void_p(), C.c_void_p(), C.c_size_t()
        L.call("mfic_run_baryon_particles" if baryons else "mfic_run_particles", self.h, C.byref(pos), C.byref(vel),
               C.byref(n))
        return pos.value, vel.value, n.value

    def particles(self, baryons: bool = False):
        """download FP64 positions and velocities, shape (npart, 3) each"""
        pos, vel, n = self.particles_device(baryons)
        out = []
        for ptr in (pos, vel):
            host = np.empty(3 * n, dtype=np.float64)
            _memcpy_d2h(host, ptr, self.ctx)
            out.append(host.reshape(n, 3))
        return out[0], out[1]

    def close(self):
        if self.h:
            L.call("mfic_run_destroy", self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _memcpy_d2h(host: np.ndarray, dptr: int, ctx: Context) -> None:
    """hipMemcpy of a device buffer into a numpy array (via the HIP runtime)."""
    hip = _hip()
    ctx.sync()
    rc = hip.hipMemcpy(C.c_void_p(host.ctypes.data), C.c_void_p(dptr), C.c_size_t(host.nbytes), 2)
    if rc != 0:
        raise RuntimeError(f"hipMemcpy failed ({rc})")


_HIP = None


def _hip():
    global _HIP
    if _HIP is None:
        for name in ("libamdhip64.so", "libamdhip64.so.7", "/opt/rocm/lib/libamdhip64.so"):
            try:
                _HIP = C.CDLL(name)
                break
            except OSError:
                continue
        if _HIP is None:
            raise RuntimeError("HIP runtime library not found")
        _HIP.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _HIP.hipMemcpy.restype = C.c_int
    return _HIP


def fft_c2c(ctx: Context, data_ptr: int, shape: Sequence[int], prec: int, forward: bool = True) -> None:
    """Grid_FFT<ccomplex_t> transform (grid_fft.cc:56-63) of a device
    [n0][n1][n2] complex array in place: unitary, forward e^{-ikx}"""
    L.call("mfic_fft_c2c", ctx.h, C.c_void_p(data_ptr), (C.c_size_t * 3)(*shape), prec, int(forward))


def ppt_fields(run: ICRun, velocity_unit: float = 1.0):
    """Semiclassical PT density and velocity of the field_eulerian output
    (ic_generator.cc:786-897) from an executed run: ([delta, vx, vy, vz] as
    real-space Grids of the run's size and precision, hbar)"""
    n, box = run.cfg.n, run.cfg.box
    outs = [Grid(run.ctx, (n,) * 3, (box,) * 3, prec=run.cfg.prec) for _ in range(4)]
    hb = C.c_double()
    L.call("mfic_ppt_fields", run.h, velocity_unit, *[g.h for g in outs], C.byref(hb))
    return outs, hb.value
