"""A minimal HDF5 reader over the HDF5 C library (ctypes), for the output
tests: datasets with their file datatype, group attributes with theirs.
Not a test module."""
import ctypes as C
import os

import numpy as np

_H = None
H5T_INTEGER, H5T_FLOAT, H5T_STRING = 0, 1, 3
H5T_SGN_NONE = 0


def lib():
    global _H
    if _H is None:
        for name in (os.environ.get("MFIC_HDF5_LIB", ""), "libhdf5.so", "libhdf5_serial.so",
                     "/opt/conda/lib/libhdf5.so"):
            if not name:
                continue
            try:
                _H = C.CDLL(name)
                break
            except OSError:
                continue
        if _H is None:
            raise OSError("HDF5 C library not found")
        h = _H
        h.H5open()
        i64 = C.c_int64
        for fn, res, args in (
            ("H5Fopen", i64, [C.c_char_p, C.c_uint, i64]), ("H5Fclose", C.c_int, [i64]),
            ("H5Dopen2", i64, [i64, C.c_char_p, i64]), ("H5Dclose", C.c_int, [i64]),
            ("H5Dget_space", i64, [i64]), ("H5Dget_type", i64, [i64]),
            ("H5Dread", C.c_int, [i64, i64, i64, i64, i64, C.c_void_p]),
            ("H5Sget_simple_extent_ndims", C.c_int, [i64]),
            ("H5Sget_simple_extent_dims", C.c_int, [i64, C.POINTER(C.c_ulonglong), C.c_void_p]),
            ("H5Sget_simple_extent_type", C.c_int, [i64]), ("H5Sclose", C.c_int, [i64]),
            ("H5Tget_class", C.c_int, [i64]), ("H5Tget_size", C.c_size_t, [i64]), ("H5Tget_sign", C.c_int, [i64]),
            ("H5Tclose", C.c_int, [i64]), ("H5Gopen2", i64, [i64, C.c_char_p, i64]), ("H5Gclose", C.c_int, [i64]),
            ("H5Aopen_by_name", i64, [i64, C.c_char_p, C.c_char_p, i64, i64]), ("H5Aclose", C.c_int, [i64]),
            ("H5Aget_type", i64, [i64]), ("H5Aget_space", i64, [i64]),
            ("H5Aread", C.c_int, [i64, i64, C.c_void_p]), ("H5Lexists", C.c_int, [i64, C.c_char_p, i64]),
            ("H5Dget_create_plist", i64, [i64]), ("H5Pclose", C.c_int, [i64]),
            ("H5Pget_layout", C.c_int, [i64]), ("H5Pget_chunk", C.c_int, [i64, C.c_int, C.POINTER(C.c_ulonglong)]),
            ("H5Pget_nfilters", C.c_int, [i64]),
            ("H5Pget_filter2", C.c_int, [i64, C.c_uint, C.POINTER(C.c_uint), C.POINTER(C.c_size_t),
                                         C.POINTER(C.c_uint), C.c_size_t, C.c_char_p, C.POINTER(C.c_uint)]),
        ):
            f = getattr(h, fn)
            f.restype, f.argtypes = res, args
    return _H


def _np_type(h, t):
    cls, size = h.H5Tget_class(t), h.H5Tget_size(t)
    if cls == H5T_FLOAT:
        return np.dtype(f"<f{size}")
    if cls == H5T_INTEGER:
        return np.dtype(("<u" if h.H5Tget_sign(t) == H5T_SGN_NONE else "<i") + str(size))
    if cls == H5T_STRING:
        return np.dtype(f"S{size}")
    raise TypeError(f"unsupported HDF5 class {cls}")


def _shape(h, sp):
    nd = h.H5Sget_simple_extent_ndims(sp)
    if nd <= 0:
        return ()
    dims = (C.c_ulonglong * nd)()
    h.H5Sget_simple_extent_dims(sp, dims, None)
    return tuple(int(d) for d in dims)


class File:
    def __init__(self, path):
        self.h = lib()
        self.f = self.h.H5Fopen(path.encode(), 0, 0)
        if self.f < 0:
            raise OSError(f"cannot open {path}")

    def close(self):
        if self.f >= 0:
            self.h.H5Fclose(self.f)
            self.f = -1

    def exists(self, path):
        return self.h.H5Lexists(self.f, path.encode(), 0) > 0

    def dataset(self, path):
        """(array in its file datatype, numpy dtype of the file datatype)"""
        h = self.h
        d = h.H5Dopen2(self.f, path.encode(), 0)
        if d < 0:
            raise KeyError(path)
        t, sp = h.H5Dget_type(d), h.H5Dget_space(d)
        dt, shape = _np_type(h, t), _shape(h, sp)
        out = np.empty(shape, dtype=dt)
        assert h.H5Dread(d, t, 0, 0, 0, out.ctypes.data) >= 0
        h.H5Tclose(t), h.H5Sclose(sp), h.H5Dclose(d)
        return out, dt

    def storage(self, path):
        """(chunk shape or None when contiguous, filter ids) of a dataset"""
        h = self.h
        d = h.H5Dopen2(self.f, path.encode(), 0)
        if d < 0:
            raise KeyError(path)
        pl = h.H5Dget_create_plist(d)
        chunk = None
        if h.H5Pget_layout(pl) == 2:  # H5D_CHUNKED
            dims = (C.c_ulonglong * 4)()
            nd = h.H5Pget_chunk(pl, 4, dims)
            chunk = tuple(int(dims[i]) for i in range(nd))
        filters = []
        for i in range(h.H5Pget_nfilters(pl)):
            flags, nel, cd = C.c_uint(), C.c_size_t(0), (C.c_uint * 1)()
            name = C.create_string_buffer(64)
            conf = C.c_uint()
            filters.append(h.H5Pget_filter2(pl, i, C.byref(flags), C.byref(nel), cd, 64, name, C.byref(conf)))
        h.H5Pclose(pl), h.H5Dclose(d)
        return chunk, filters

    def attr(self, group, name):
        """(value, numpy dtype, shape) of a group attribute; scalars have
        shape ()"""
        h = self.h
        a = h.H5Aopen_by_name(self.f, group.encode(), name.encode(), 0, 0)
        if a < 0:
            raise KeyError(f"{group}:{name}")
        t, sp = h.H5Aget_type(a), h.H5Aget_space(a)
        dt, shape = _np_type(h, t), _shape(h, sp)
        out = np.empty(shape if shape else (1,), dtype=dt)
        assert h.H5Aread(a, t, out.ctypes.data) >= 0
        h.H5Tclose(t), h.H5Sclose(sp), h.H5Aclose(a)
        return (out if shape else out[0]), dt, shape
