"""Host-side mirror of the reference's ``Grid_FFT`` on top of the mfic C ABI.

``Context`` owns one HIP device and stream; ``Grid`` owns one device buffer in
the reference layout (grid_fft.cc:40-112): ``n0*n1*(n2+2)`` reals, k-space view
complex ``[n0][n1][n2/2+1]``.  Method names follow ``Grid_FFT`` (grid_fft.hh).
All work is enqueued on the context stream; ``download`` synchronises.
"""
from __future__ import annotations

import ctypes as C
import weakref
from typing import Sequence

import numpy as np

from . import _lib as L


class Context:
    """One HIP device + stream + twiddle caches (mfic_ctx)."""

    def __init__(self, device: int = 0):
        self.h = C.c_void_p()
        L.call("mfic_ctx_create", device, C.byref(self.h))
        self.device = device
        self._children = weakref.WeakSet()

    def adopt(self, obj) -> None:
        """register a handle owner (grid, run, ...) to be closed before the context"""
        self._children.add(obj)

    def sync(self) -> None:
        L.call("mfic_ctx_sync", self.h)

    @property
    def stream(self) -> int:
        s = C.c_void_p()
        L.call("mfic_ctx_stream", self.h, C.byref(s))
        return s.value or 0

    def close(self) -> None:
        if self.h:
            for child in list(self._children):
                child.close()
            L.call("mfic_ctx_destroy", self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _np_dtype(prec: int):
    return np.float64 if prec == L.F64 else np.float32


class Grid:
    """Grid_FFT<real_t> equivalent living in device memory."""

    def __init__(self, ctx: Context, n: Sequence[int], length: Sequence[float], prec: int = L.F32,
                 space: int = L.RSPACE, handle=None):
        self.ctx = ctx
        self.n = tuple(int(x) for x in (n if len(n) == 3 else (n[0],) * 3))
        self.length = tuple(float(x) for x in (length if len(length) == 3 else (length[0],) * 3))
        self.prec = prec
        self.owned = handle is None
        if handle is None:
            self.h = C.c_void_p()
            L.call("mfic_grid_create", ctx.h, (C.c_size_t * 3)(*self.n), (C.c_double * 3)(*self.length),
                   prec, space, C.byref(self.h))
            ctx.adopt(self)
        else:
            self.h = handle

    # -- storage --------------------------------------------------------------
    @property
    def shape_real(self):
        return (self.n[0], self.n[1], self.n[2] + 2)

    @property
    def nc(self) -> int:
        return self.n[2] // 2 + 1

    @property
    def space(self) -> int:
        s = C.c_int()
        L.call("mfic_grid_get_space", self.h, C.byref(s))
        return s.value

    def set_space(self, space: int) -> None:
        L.call("mfic_grid_set_space", self.h, space)

    def device_ptr(self) -> int:
        p, n = C.c_void_p(), C.c_size_t()
        L.call("mfic_grid_device_ptr", self.h, C.byref(p), C.byref(n))
        return p.value

    def upload(self, a: np.ndarray) -> None:
        """copy a (n0, n1, n2+2) real array (float32/float64) to the device"""
        a = np.ascontiguousarray(a)
        if a.shape != self.shape_real:
            raise ValueError(f"expected shape {self.shape_real}, got {a.shape}")
        hp = L.F64 if a.dtype == np.float64 else L.F32
        if a.dtype not in (np.float32, np.float64):
            raise TypeError("float32 or float64 arrays only")
        L.call("mfic_grid_upload", self.h, a.ctypes.data, hp)

    def upload_k(self, k: np.ndarray) -> None:
        """copy a complex (n0, n1, n2/2+1) array and mark the grid k-space"""
        k = np.ascontiguousarray(k, dtype=np.complex128)
        self.upload(k.view(np.float64).reshape(self.shape_real))
        self.set_space(L.KSPACE)

    def download(self, dtype=None) -> np.ndarray:
        dtype = dtype or _np_dtype(self.prec)
        out = np.empty(self.shape_real, dtype=dtype)
        L.call("mfic_grid_download", self.h, out.ctypes.data, L.F64 if out.dtype == np.float64 else L.F32)
        return out

    def download_k(self) -> np.ndarray:
        return self.download(np.float64).view(np.complex128)

    def download_r(self) -> np.ndarray:
        return self.download(np.float64)[:, :, : self.n[2]]

    # -- Grid_FFT operations --------------------------------------------------
    def FourierTransformForward(self, do_transform: bool = True) -> None:
        if do_transform:
            L.call("mfic_fft_forward", self.h)
        else:
            self.set_space(L.KSPACE)

    def FourierTransformBackward(self, do_transform: bool = True) -> None:
        if do_transform:
            L.call("mfic_fft_backward", self.h)
        else:
            self.set_space(L.RSPACE)

    def zero(self) -> None:
        L.call("mfic_grid_zero", self.h)

    def zero_DC_mode(self) -> None:
        L.call("mfic_grid_zero_dc", self.h)

    def apply_InverseLaplacian(self) -> None:
        L.call("mfic_grid_inverse_laplacian", self.h)

    def shift_field(self, s: Sequence[float], transform_back: bool = True) -> None:
        L.call("mfic_grid_shift", self.h, (C.c_double * 3)(*s), int(transform_back))

    def copy_from(self, other: "Grid") -> None:
        L.call("mfic_grid_copy", self.h, other.h)

    def __imul__(self, x: float) -> "Grid":
        L.call("mfic_grid_scale", self.h, float(x))
        return self

    def close(self) -> None:
        if self.owned and self.h:
            L.call("mfic_grid_destroy", self.h)
        self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
