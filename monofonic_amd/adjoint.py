"""Differentiable 2LPT: the white noise -> displacement map of the LPT block
(ic_generator.cc:378-383, 509-555, 931-970) and its hand-written adjoint
(``mfic_lpt2_forward`` / ``mfic_lpt2_adjoint``), exposed to PyTorch as an
``autograd.Function``.  PyTorch supplies device memory and the loss around
the map; the map and its reverse mode run in the HIP library.  The adjoint
recomputes phi1 from the noise rather than keeping forward intermediates.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

from . import _lib as L
from .grid import Context, Grid
from .lpt import baseline_cosmology


def _torch():
    import torch
    return torch


class LPT2:
    """Plan of the 2LPT map on an n^3 grid of side `box` (mfic_lpt2).
    Displacements come out in Mpc/h times `unit`."""

    def __init__(self, ctx: Context, n: int, box: float, prec: int = L.F32, zstart: float = 24.0,
                 cosmo: Optional[L.CosmoParams] = None, unit: float = 1.0):
        self.ctx, self.n, self.box, self.prec = ctx, n, box, prec
        self.cosmo = cosmo if cosmo is not None else baseline_cosmology(zstart)
        self.cosmo.zstart = zstart
        self.h = C.c_void_p()
        L.call("mfic_lpt2_create", ctx.h, n, box, prec, C.byref(self.cosmo), unit, C.byref(self.h))
        ctx.adopt(self)

    # -- grid level ----------------------------------------------------------
    def forward_grids(self, noise: Grid, psi: Sequence[Grid]) -> None:
        L.call("mfic_lpt2_forward", self.h, noise.h, *[g.h for g in psi])

    def adjoint_grids(self, noise: Grid, cot: Sequence[Grid], grad: Grid) -> None:
        L.call("mfic_lpt2_adjoint", self.h, noise.h, *[g.h for g in cot], grad.h)

    # -- torch tensors -------------------------------------------------------
    def _dtype(self):
        torch = _torch()
        return torch.float64 if self.prec == L.F64 else torch.float32

    def _wrap(self, t) -> Grid:
        """grid over a contiguous torch tensor of shape (n, n, n + 2)"""
        n = self.n
        assert t.is_contiguous() and tuple(t.shape) == (n, n, n + 2) and t.dtype == self._dtype()
        h = C.c_void_p()
        L.call("mfic_grid_wrap", self.ctx.h, (C.c_size_t * 3)(n, n, n), (C.c_double * 3)(*(self.box,) * 3),
               self.prec, L.RSPACE, C.c_void_p(t.data_ptr()), C.byref(h))
        return Grid(self.ctx, (n,) * 3, (self.box,) * 3, prec=self.prec, handle=h)

    def _padded(self, x):
        torch = _torch()
        n = self.n
        buf = torch.zeros((n, n, n + 2), dtype=self._dtype(), device=x.device)
        buf[:, :, :n] = x
        return buf

    def _call(self, fn, tensors):
        torch = _torch()
        grids = [self._wrap(t) for t in tensors]
        torch.cuda.synchronize()
        try:
            fn(grids)
            self.ctx.sync()
        finally:
            for g in grids:
                L.call("mfic_grid_destroy", g.h)
                g.h = C.c_void_p()

    def forward(self, noise):
        """noise (n, n, n) -> displacement (3, n, n, n)"""
        torch = _torch()
        n = self.n
        src = self._padded(noise.detach())
        out = torch.empty((3, n, n, n + 2), dtype=self._dtype(), device=noise.device)
        self._call(lambda g: self.forward_grids(g[0], g[1:]), [src] + [out[d] for d in range(3)])
        return out[:, :, :, :n]

    def adjoint(self, noise, cot):
        """noise (n, n, n), cotangent (3, n, n, n) -> d(loss)/d(noise) (n, n, n)"""
        torch = _torch()
        n = self.n
        src = self._padded(noise.detach())
        cots = [self._padded(cot[d]) for d in range(3)]
        grad = torch.empty((n, n, n + 2), dtype=self._dtype(), device=noise.device)
        self._call(lambda g: self.adjoint_grids(g[0], g[1:4], g[4]), [src] + cots + [grad])
        return grad[:, :, :n]

    def close(self) -> None:
        if self.h:
            L.call("mfic_lpt2_destroy", self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _make_function():
    torch = _torch()

    class LPT2Map(torch.autograd.Function):
        """psi = LPT2Map.apply(noise, plan): (n, n, n) -> (3, n, n, n)"""

        @staticmethod
        def forward(ctx, noise, plan):
            ctx.plan = plan
            ctx.save_for_backward(noise)
            return plan.forward(noise)

        @staticmethod
        def backward(ctx, gpsi):
            (noise,) = ctx.saved_tensors
            return ctx.plan.adjoint(noise, gpsi.contiguous()), None

    return LPT2Map


_FN = None


def lpt2_map(noise, plan: LPT2):
    """differentiable 2LPT displacement of `noise` under `plan`"""
    global _FN
    if _FN is None:
        _FN = _make_function()
    return _FN.apply(noise, plan)


def cic_density(psi, box: float):
    """cloud-in-cell density (particles per cell, mean 1) on the n^3 mesh of
    one particle per cell displaced by psi (3, n, n, n) in Mpc/h;
    differentiable in psi through torch autograd.  The cell index and the
    CIC weights come from the displacement alone (in cells), and the
    particle's integer lattice index is added afterwards, so the weights keep
    the displacement's own precision instead of that of q + psi."""
    torch = _torch()
    n = psi.shape[1]
    ar = torch.arange(n, device=psi.device, dtype=torch.long)
    q = torch.stack(torch.meshgrid(ar, ar, ar, indexing="ij"))
    s = psi * (n / box)
    i0 = torch.floor(s)
    f = s - i0
    i0 = i0.long() + q
    rho = torch.zeros(n * n * n, dtype=psi.dtype, device=psi.device)
    for dx in (0, 1):
        wx = f[0] if dx else 1 - f[0]
        ix = (i0[0] + dx) % n
        for dy in (0, 1):
            wy = f[1] if dy else 1 - f[1]
            iy = (i0[1] + dy) % n
            for dz in (0, 1):
                wz = f[2] if dz else 1 - f[2]
                iz = (i0[2] + dz) % n
                rho = rho.index_add(0, ((ix * n + iy) * n + iz).reshape(-1), (wx * wy * wz).reshape(-1))
    return rho.reshape(n, n, n)
