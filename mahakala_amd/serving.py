"""Continuous batching of image and shadow requests on one GPU.

Python face of the ``mk_server_*`` entry points.  The reference answers one
``make_image`` (images.py:30-144) or ``find_shadow_bisection_angles``
(geodesics.py:405-435) call at a time; a ``Server`` accepts many such
requests, of different resolution, spin, inclination and frequency, and runs
their rays together in shared launches on resident snapshots.  Each request's
result is the same as when it runs alone.

    srv = Server(precision="mixed")
    sid = srv.add_snapshot(model)                    # AthenakFluidModel
    rid = srv.submit_image(sid, bhspin=0.5, inclination=30, resolution=128)
    img = srv.result(rid, timeout=60)                # (128, 128) FP64
"""
from __future__ import annotations

import ctypes as C
import threading

import numpy as np

from . import _lib
from .constants import Msun

PRECISIONS = {"plain64": 0, "fp64": 0, "mixed": 1}


class Server:
    """Owner of an ``mk_server`` (destroyed by ``close`` or on collection)."""

    def __init__(self, precision="plain64", lanes=0, steps_per_launch=0, max_requests=0, hbm_budget=0,
                 use_graph=True, persistent=True, shadow_in_launch=False, image_in_launch=False):
        """persistent: one resident launch whose waves take newly admitted
        work without a launch boundary (MK_SERVER_PERSISTENT; lower latency
        and higher throughput on every stream measured); False: a launch per
        steps_per_launch steps (use_graph applies to this form only).
        shadow_in_launch: bisections always run inside the launches
        (MK_SERVER_SHADOW_IN_LAUNCH); by default an idle server runs a
        bisection of up to 1024 angles itself, in the band schedule of
        find_shadow_bisection_angles (same radii, about a tenth of the
        latency).  image_in_launch: images always run inside the launches
        (MK_SERVER_IMAGE_IN_LAUNCH); by default an idle server whose queue
        holds one image of up to 512^2 and nothing else renders it itself
        through mk_render (the same image)."""
        if precision not in PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(PRECISIONS)}")
        flags = (1 if persistent else 0) | (2 if shadow_in_launch else 0) | (4 if image_in_launch else 0)
        cfg = _lib.mk_server_config(PRECISIONS[precision], int(lanes), int(steps_per_launch), int(max_requests),
                                    int(hbm_budget), int(bool(use_graph)), flags)
        h = _lib.vp()
        _lib.check(_lib.lib().mk_server_create(C.byref(cfg), C.byref(h)))
        self._h = h.value
        self._snaps = []          # snapshot owners stay alive with the server
        self._out = {}            # request id -> (output array, shape)
        self._done = {}           # request id -> latency (ms)
        self._mu = threading.Lock()

    # ------------------------------------------------------------ requests
    def add_snapshot(self, model_or_snapshot):
        """Registers a resident snapshot (an AthenakFluidModel or its
        device snapshot) and returns its id."""
        snap = getattr(model_or_snapshot, "device_snapshot", model_or_snapshot)
        sid = C.c_int32()
        _lib.check(_lib.lib().mk_server_add_snapshot(self._h, snap.handle, C.byref(sid)))
        self._snaps.append(snap)
        return sid.value

    def submit_image(self, snapshot, bhspin, inclination=60, distance=1000, fov=20, resolution=160,
                     observing_frequency=230.e9, mass_scale=1.e26, M_bh=6.2e9 * Msun, r_high=40,
                     fluid_gamma=13 / 9, max_nsteps=10000, priority=0):
        """Queues an image request (make_image's parameters); returns its id.
        A higher priority is admitted before queued requests of lower
        priority; equal priorities keep submit order."""
        res = int(resolution)
        req = _lib.mk_image_request(int(snapshot), res, int(max_nsteps), int(priority), float(bhspin),
                                    float(inclination),
                                    float(distance), float(fov), float(observing_frequency), float(mass_scale),
                                    float(M_bh), float(r_high), float(fluid_gamma))
        out = np.empty((res, res), dtype=np.float64)
        rid = C.c_int64()
        with self._mu:
            _lib.check(_lib.lib().mk_server_submit_image(self._h, C.byref(req), out.ctypes.data, C.byref(rid)))
            self._out[rid.value] = out
        return rid.value

    def submit_shadow(self, bhspin, inclination, angles, distance=1000.0, max_steps=2000, error_allowed=1e-3,
                      max_it=40, priority=0):
        """Queues a shadow bisection (find_shadow_bisection_angles); returns
        its id.  The result is the radius per angle.  priority: as for
        images."""
        ang = np.ascontiguousarray(angles, dtype=np.float64).ravel()
        req = _lib.mk_shadow_request(float(bhspin), float(inclination), float(distance), float(error_allowed),
                                     ang.size, int(max_steps), int(max_it), int(priority), ang.ctypes.data)
        out = np.empty(ang.size, dtype=np.float64)
        rid = C.c_int64()
        with self._mu:
            _lib.check(_lib.lib().mk_server_submit_shadow(self._h, C.byref(req), out.ctypes.data, C.byref(rid)))
            self._out[rid.value] = out
        return rid.value

    # ---------------------------------------------------------- completion
    def poll(self, cap=4096):
        """Completed requests since the last poll: list of (id, latency_ms)."""
        ids = np.empty(cap, dtype=np.int64)
        lat = np.empty(cap, dtype=np.float64)
        n = C.c_int32()
        _lib.check(_lib.lib().mk_server_poll(self._h, ids.ctypes.data, lat.ctypes.data, cap, C.byref(n)))
        got = list(zip(ids[:n.value].tolist(), lat[:n.value].tolist()))
        with self._mu:
            self._done.update(got)
        return got

    def wait(self, rid=-1, timeout=600.0):
        """Blocks until request ``rid`` (or every request, rid < 0) is done."""
        _lib.check(_lib.lib().mk_server_wait(self._h, int(rid), float(timeout)))

    def result(self, rid, timeout=600.0):
        """Waits for ``rid`` and returns its output (image or radii)."""
        self.wait(rid, timeout)
        with self._mu:
            return self._out.pop(rid)

    def cancel(self, rid):
        """Removes request ``rid`` if it is still queued: True, and a later
        result() raises MkError; False when it already runs or is done."""
        done = C.c_int32()
        _lib.check(_lib.lib().mk_server_cancel(self._h, int(rid), C.byref(done)))
        if done.value:
            with self._mu:
                self._out.pop(rid, None)
        return bool(done.value)

    def stats(self):
        st = _lib.mk_server_stats()
        _lib.check(_lib.lib().mk_server_get_stats(self._h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in st._fields_}

    def close(self):
        if getattr(self, "_h", None):
            _lib.check(_lib.lib().mk_server_destroy(self._h))
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        if getattr(self, "_h", None) and _lib._lib is not None:
            _lib._lib.mk_server_destroy(self._h)
            self._h = None
