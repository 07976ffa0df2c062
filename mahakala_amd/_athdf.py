"""ctypes binding of libmk_athdf.so, the native AthenaK .athdf reader.

It stands in for the h5py reads of athenak.py:79-101 where h5py is not
installed: the ten datasets come back as NumPy arrays in their stored dtype
and VariableNames as a list of str.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libmk_athdf.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "mk_athdf.h")
DATASETS = ("x1v", "x2v", "x3v", "x1f", "x2f", "x3f", "uov", "B", "LogicalLocations", "Levels")
_TYPES = {1: np.float32, 2: np.float64, 3: np.int32, 4: np.int64}

_lib = None


class AthdfError(OSError):
    """A libmk_athdf call failed."""


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise AthdfError(f"libmk_athdf.so not built at {LIB_PATH} (needs the HDF5 C library)")
        L = C.CDLL(LIB_PATH)
        vp, i32p = C.c_void_p, C.POINTER(C.c_int32)
        L.mk_athdf_last_error.restype = C.c_char_p
        L.mk_athdf_last_error.argtypes = []
        L.mk_athdf_open.argtypes = [C.c_char_p, C.POINTER(vp)]
        L.mk_athdf_close.argtypes = [vp]
        L.mk_athdf_info.argtypes = [vp, C.c_char_p, i32p, i32p, C.POINTER(C.c_int64)]
        L.mk_athdf_read.argtypes = [vp, C.c_char_p, C.c_int32, vp]
        L.mk_athdf_varnames.argtypes = [vp, C.c_char_p, C.c_size_t, i32p]
        for n in ("mk_athdf_open", "mk_athdf_close", "mk_athdf_info", "mk_athdf_read", "mk_athdf_varnames"):
            getattr(L, n).restype = C.c_int
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise AthdfError(f"libmk_athdf error {rc}: {lib().mk_athdf_last_error().decode()}")


def read_athdf_native(filename):
    """Datasets and VariableNames of an .athdf file, as h5py would give them."""
    L = lib()
    h = C.c_void_p()
    _check(L.mk_athdf_open(os.fsencode(filename), C.byref(h)))
    try:
        out = {}
        for name in DATASETS:
            t, nd = C.c_int32(), C.c_int32()
            dims = (C.c_int64 * 8)()
            _check(L.mk_athdf_info(h, name.encode(), C.byref(t), C.byref(nd), dims))
            arr = np.empty(tuple(dims[:nd.value]), dtype=_TYPES[t.value])
            _check(L.mk_athdf_read(h, name.encode(), t.value, arr.ctypes.data))
            out[name] = arr
        buf = C.create_string_buffer(1 << 16)
        n = C.c_int32()
        _check(L.mk_athdf_varnames(h, buf, len(buf), C.byref(n)))
        out["VariableNames"] = buf.raw.split(b"\0")[:n.value]
        out["VariableNames"] = [v.decode("utf-8") for v in out["VariableNames"]]
        return out
    finally:
        L.mk_athdf_close(h)
