"""ctypes binding of libmahakala_hip.so (the HIP kernels for gfx950).

The library is built in-tree by ``__graft_entry__.build()`` (``make -C
mahakala_amd/csrc``) into ``mahakala_amd/_lib/libmahakala_hip.so``.  There
is no CPU fallback: if the library or a GPU is missing, calls raise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libmahakala_hip.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "mahakala_hip.h")

_lib = None

vp = C.c_void_p
dbl = C.c_double
i32 = C.c_int32
lng = C.c_long


class MkError(RuntimeError):
    """A libmahakala_hip call returned a negative status."""


class mk_render_params(C.Structure):
    _fields_ = [
        ("bhspin", dbl), ("inclination", dbl), ("distance", dbl), ("fov", dbl),
        ("mass_scale", dbl), ("M_bh", dbl), ("r_high", dbl), ("fluid_gamma", dbl),
        ("observing_frequency", dbl), ("div", dbl), ("tol", dbl), ("fallback_pitch", dbl),
        ("sigma_cut", dbl), ("fp64_radius", dbl), ("ring_radius", dbl), ("ring_drdl", dbl),
        ("resolution", i32), ("max_nsteps", i32), ("precision", i32), ("flags", i32),
    ]


class mk_server_config(C.Structure):
    _fields_ = [
        ("precision", i32), ("lanes", i32), ("steps_per_launch", i32), ("max_requests", i32),
        ("hbm_budget", C.c_size_t), ("use_graph", i32), ("flags", i32),
    ]


class mk_image_request(C.Structure):
    _fields_ = [
        ("snapshot", i32), ("resolution", i32), ("max_nsteps", i32), ("priority", i32),
        ("bhspin", dbl), ("inclination", dbl), ("distance", dbl), ("fov", dbl),
        ("observing_frequency", dbl), ("mass_scale", dbl), ("M_bh", dbl), ("r_high", dbl),
        ("fluid_gamma", dbl),
    ]


class mk_shadow_request(C.Structure):
    _fields_ = [
        ("bhspin", dbl), ("inclination", dbl), ("distance", dbl), ("error_allowed", dbl),
        ("n_angles", i32), ("max_steps", i32), ("max_it", i32), ("priority", i32),
        ("h_angles", vp),
    ]


class mk_server_stats(C.Structure):
    _fields_ = [
        ("submitted", C.c_int64), ("admitted", C.c_int64), ("completed", C.c_int64),
        ("queued", C.c_int64), ("in_flight", C.c_int64), ("items_done", C.c_int64),
        ("iterations", C.c_int64), ("gpu_ms", dbl), ("hbm_in_use", C.c_size_t),
        ("hbm_budget", C.c_size_t), ("lane_steps_busy", C.c_int64), ("lane_steps_total", C.c_int64),
        ("lanes", i32), ("steps_per_launch", i32),
    ]


# name -> argtypes (restype int unless listed in _RESTYPES)
_SIGS = {
    "mk_last_error": [],
    "mk_version": [],
    "mk_device_info": [C.c_char_p, C.c_size_t, C.POINTER(C.c_int), C.POINTER(C.c_size_t)],
    "mk_device_mem": [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)],
    "mk_malloc": [C.POINTER(vp), C.c_size_t],
    "mk_free": [vp],
    "mk_memcpy_h2d": [vp, vp, C.c_size_t, vp],
    "mk_memcpy_d2h": [vp, vp, C.c_size_t, vp],
    "mk_memset": [vp, C.c_int, C.c_size_t, vp],
    "mk_device_copy": [vp, vp, C.c_size_t, vp],
    "mk_step_cycles": [C.c_int, C.c_int, dbl, vp],
    "mk_stream_create": [C.POINTER(vp)],
    "mk_stream_destroy": [vp],
    "mk_stream_sync": [vp],
    "mk_event_create": [C.POINTER(vp)],
    "mk_event_destroy": [vp],
    "mk_event_record": [vp, vp],
    "mk_event_sync": [vp],
    "mk_event_elapsed_ms": [vp, vp, C.POINTER(C.c_float)],
    "mk_camera_grid": [dbl, dbl, dbl, dbl, dbl, C.c_int, vp, vp],
    "mk_camera_pixels": [dbl, dbl, dbl, vp, vp, lng, vp, vp],
    "mk_initial_condition": [vp, vp, lng, dbl, vp, vp],
    "mk_metric": [vp, lng, dbl, vp, vp],
    "mk_imetric": [vp, lng, dbl, vp, vp],
    "mk_rhs": [vp, lng, dbl, vp, vp],
    "mk_rk4": [vp, vp, lng, dbl, vp, vp],
    "mk_radius_cal": [vp, lng, C.c_int, dbl, vp, vp],
    "mk_integrate_final": [vp, lng, C.c_int, dbl, dbl, dbl, vp, vp, vp, vp],
    "mk_trim_rows": [vp, lng, C.c_int, C.POINTER(lng), vp],
    "mk_integrate_history": [vp, lng, C.c_int, dbl, dbl, dbl, lng, vp, vp, vp],
    "mk_select_photons": [dbl, vp, vp, lng, dbl, dbl, C.c_int, vp],
    "mk_shadow_bisect": [dbl, dbl, vp, lng, C.c_int, dbl, C.c_int, dbl, vp],
    "mk_shadow_bisect_noalloc": [dbl, dbl, vp, lng, C.c_int, dbl, C.c_int, dbl, vp],
    "mk_synchrotron": [vp, vp, vp, vp, vp, lng, C.c_int, dbl, vp, vp, vp],
    "mk_solve_intensity": [vp, vp, vp, lng, lng, dbl, vp, vp, vp],
    "mk_attenuated_emissivity": [vp, vp, vp, lng, lng, dbl, vp, vp],
    "mk_rlow_rhigh": [vp, vp, vp, lng, dbl, dbl, dbl, dbl, vp, vp],
    "mk_snapshot_create": [vp, vp, lng, C.c_int, vp, vp, vp, vp, vp, vp, vp, vp, C.POINTER(vp)],
    "mk_snapshot_free": [vp],
    "mk_snapshot_bytes": [vp, C.POINTER(C.c_size_t)],
    "mk_snapshot_download": [vp, vp],
    "mk_block_index": [vp, vp, lng, C.c_int, vp, vp],
    "mk_sample_prims": [vp, vp, lng, vp, vp],
    "mk_sample_scalars": [vp, vp, lng, dbl, dbl, vp, vp],
    "mk_render": [vp, C.POINTER(mk_render_params), vp, vp, vp],
    "mk_render_form": [vp, C.POINTER(mk_render_params), C.POINTER(i32)],
    "mk_server_create": [C.POINTER(mk_server_config), C.POINTER(vp)],
    "mk_server_add_snapshot": [vp, vp, C.POINTER(i32)],
    "mk_server_submit_image": [vp, C.POINTER(mk_image_request), vp, C.POINTER(C.c_int64)],
    "mk_server_submit_shadow": [vp, C.POINTER(mk_shadow_request), vp, C.POINTER(C.c_int64)],
    "mk_server_poll": [vp, vp, vp, i32, C.POINTER(i32)],
    "mk_server_wait": [vp, C.c_int64, dbl],
    "mk_server_cancel": [vp, C.c_int64, C.POINTER(i32)],
    "mk_server_get_stats": [vp, C.POINTER(mk_server_stats)],
    "mk_server_destroy": [vp],
}
_RESTYPES = {"mk_last_error": C.c_char_p}


def lib():
    """Load the HIP library (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MkError(f"libmahakala_hip.so not built at {LIB_PATH}; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = _RESTYPES.get(name, C.c_int)
        _lib = L
    return _lib


def declared_symbols():
    """Names of the entry points declared in include/mahakala_hip.h."""
    import re
    text = open(HEADER_PATH).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\*?(mk_\w+)\s*\(", text, re.M)))


def check(rc):
    if rc != 0:
        msg = lib().mk_last_error()
        raise MkError(f"libmahakala_hip error {rc}: {msg.decode() if msg else ''}")
    return rc


class DeviceBuffer:
    """Device allocation owned by Python (freed on garbage collection)."""

    __slots__ = ("ptr", "nbytes")

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = vp()
        check(lib().mk_malloc(C.byref(p), max(self.nbytes, 1)))
        self.ptr = p.value

    @classmethod
    def from_array(cls, arr, dtype=np.float64):
        a = np.ascontiguousarray(arr, dtype=dtype)
        buf = cls(a.nbytes)
        if a.nbytes:
            check(lib().mk_memcpy_h2d(buf.ptr, a.ctypes.data, a.nbytes, None))
        return buf

    def to_array(self, shape, dtype=np.float64):
        out = np.empty(shape, dtype=dtype)
        if out.nbytes:
            check(lib().mk_memcpy_d2h(out.ctypes.data, self.ptr, out.nbytes, None))
        return out

    def __del__(self):
        if getattr(self, "ptr", None) and _lib is not None:
            _lib.mk_free(self.ptr)
            self.ptr = None


def device_memory():
    """(free, total) bytes of the current device."""
    free, total = C.c_size_t(), C.c_size_t()
    check(lib().mk_device_mem(C.byref(free), C.byref(total)))
    return free.value, total.value


def device_info():
    name = C.create_string_buffer(128)
    cus = C.c_int()
    total = C.c_size_t()
    check(lib().mk_device_info(name, 128, C.byref(cus), C.byref(total)))
    return name.value.decode(), cus.value, total.value


def sync():
    check(lib().mk_stream_sync(None))
