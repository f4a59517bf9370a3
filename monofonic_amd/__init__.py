"""monofonic_amd: MI355X-native monofonIC initial-condition hot path.

The compute path is the HIP/gfx950 library ``libmfic.so`` (C ABI in
``include/mfic.h``); this package holds its ctypes binding and a host-side
mirror of the reference's Grid_FFT / OrszagConvolver / ic_generator
interfaces.  There is no CPU fallback for any device operation.
"""
from . import _lib
from ._lib import BCC, F32, F64, KSPACE, RSPACE, SC, MficError
from .grid import Context, Grid

__all__ = ["Context", "Grid", "F32", "F64", "KSPACE", "RSPACE", "SC", "BCC", "MficError", "_lib"]
