"""ctypes wrapper of oracle/_build/libmk_oracle.so (test infrastructure only).

Built by ``make -C oracle``.  Functions mirror ``mahakala_np`` where both
exist; see ``c/mk_oracle.c`` for the reference file:line each one follows.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libmk_oracle.so")
_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.mko_metric.argtypes = [_dp, C.c_double, _dp]
        L.mko_rhs.argtypes = [_dp, C.c_double, _dp, C.c_long]
        L.mko_rk4.argtypes = [_dp, _dp, C.c_double, _dp, C.c_long]
        L.mko_camera_grid.argtypes = [C.c_double] * 5 + [C.c_int, _dp]
        L.mko_camera_pixels.argtypes = [C.c_double] * 3 + [_dp, _dp, C.c_long, _dp]
        L.mko_integrate_final.argtypes = [_dp, C.c_long, C.c_int, C.c_double, C.c_double, C.c_double,
                                          _ip, _dp, _dp, C.c_void_p, C.c_int]
        L.mko_integrate_history.argtypes = [_dp, C.c_long, C.c_int, C.c_double, C.c_double, C.c_double,
                                            C.c_long, _dp, _dp, C.c_int]
        L.mko_bisect_iterations.argtypes = [C.c_double, C.c_int]
        L.mko_bisect_iterations.restype = C.c_int
        L.mko_shadow_bisect.argtypes = [C.c_double, C.c_double, _dp, C.c_long, C.c_int, C.c_double,
                                        C.c_int, C.c_double, _dp, C.c_int]
        L.mko_sample_scalars.argtypes = [_fp, C.c_long, C.c_int, _dp, _dp, _dp, _dp, C.c_void_p, C.c_int,
                                         C.c_void_p, C.c_double, _dp, C.c_long, C.c_double, C.c_double,
                                         _dp, C.c_void_p, C.c_int]
        L.mko_synchrotron.argtypes = [_dp] * 5 + [C.c_long, C.c_int, C.c_double, _dp, _dp]
        L.mko_render.argtypes = ([_fp, C.c_long, C.c_int, _dp, _dp, _dp, _dp, _ip, C.c_int, _dp, C.c_double]
                                 + [C.c_double] * 4 + [C.c_int, C.c_int] + [C.c_double] * 5
                                 + [C.c_long, C.c_long, C.c_long, _dp, _ip, C.c_void_p, C.c_void_p, C.c_int])
        L.mko_max_threads.restype = C.c_int
        _lib = L
    return _lib


def threads(n=None):
    return int(n) if n else lib().mko_max_threads()


def metric(x, a):
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(4)
    g = np.empty(16)
    lib().mko_metric(x, a, g)
    return g.reshape(4, 4)


def rhs(s, a):
    s = np.ascontiguousarray(s, dtype=np.float64)
    out = np.empty_like(s)
    lib().mko_rhs(s.reshape(-1), a, out.reshape(-1), s.size // 8)
    return out


def rk4(s, dt, a):
    s = np.ascontiguousarray(s, dtype=np.float64)
    dt = np.ascontiguousarray(np.broadcast_to(dt, s.shape[:-1]), dtype=np.float64)
    out = np.empty_like(s)
    lib().mko_rk4(s.reshape(-1), dt.reshape(-1), a, out.reshape(-1), s.size // 8)
    return out


def camera_grid(a, inc, dist, lo, hi, n):
    out = np.empty((n * n, 8))
    lib().mko_camera_grid(a, inc, dist, lo, hi, n, out)
    return out


def camera_pixels(a, inc, dist, radius, angle):
    radius = np.ascontiguousarray(radius, dtype=np.float64)
    angle = np.ascontiguousarray(np.broadcast_to(angle, radius.shape), dtype=np.float64)
    out = np.empty((radius.size, 8))
    lib().mko_camera_pixels(a, inc, dist, radius, angle, radius.size, out)
    return out


def integrate_final(s0, N, div, tol, a, nthreads=None, margins=False):
    """-> (K, last state, select value) per ray, and with margins=True the
    (n, 2) threshold margins of the rejected candidate and the last accepted
    state (see CSnapshot.render)."""
    s0 = np.ascontiguousarray(s0, dtype=np.float64)
    n = s0.shape[0]
    K = np.empty(n, dtype=np.int32)
    fin = np.empty((n, 8))
    sel = np.empty(n)
    mg = np.empty((n, 2)) if margins else None
    lib().mko_integrate_final(s0, n, N, div, tol, a, K, fin, sel,
                              None if mg is None else mg.ctypes.data_as(C.c_void_p), threads(nthreads))
    return (K, fin, sel, mg) if margins else (K, fin, sel)


def trim_rows(K, N):
    """T of the reference's trim rule from per-ray first frozen rows."""
    K = np.asarray(K)
    if K.size == 0:
        return N
    if np.any(K >= N):
        return N
    m = int(K.max())
    return N if m < 1 else min(m + 2, N)


def geodesic_integrator(N, s0, div, tol, a, nthreads=None):
    s0 = np.ascontiguousarray(s0, dtype=np.float64)
    K, _, _ = integrate_final(s0, N, div, tol, a, nthreads)
    T = trim_rows(K, N)
    S = np.empty((T, s0.shape[0], 8))
    D = np.empty((T, s0.shape[0]))
    lib().mko_integrate_history(s0, s0.shape[0], N, div, tol, a, T, S, D, threads(nthreads))
    return S, D


def bisect_iterations(error_allowed=1e-3, max_it=40):
    return lib().mko_bisect_iterations(error_allowed, max_it)


def find_shadow_bisection_angles(a, inc, angles, max_steps=2000, error_allowed=1e-3, max_it=40,
                                 distance=1000.0, nthreads=None):
    angles = np.ascontiguousarray(angles, dtype=np.float64)
    out = np.empty_like(angles)
    lib().mko_shadow_bisect(a, inc, angles, angles.size, max_steps, error_allowed, max_it, distance, out,
                            threads(nthreads))
    return out


class CSnapshot:
    """Snapshot arrays in the layout the C oracle reads."""

    def __init__(self, blocks_canon, lo, hi, v0, dx, grid=None, g0=None, gh=0.0):
        self.data = np.ascontiguousarray(blocks_canon, dtype=np.float32)
        self.nmb = self.data.shape[0]
        self.n = self.data.shape[-1] - 2
        self.lo, self.hi, self.v0, self.dx = (np.ascontiguousarray(v, dtype=np.float64) for v in (lo, hi, v0, dx))
        self.grid = None if grid is None else np.ascontiguousarray(grid, dtype=np.int32)
        self.G = 0 if grid is None else self.grid.shape[0]
        self.g0 = np.zeros(3) if g0 is None else np.ascontiguousarray(g0, dtype=np.float64)
        self.gh = float(gh)

    def _args(self):
        gp = None if self.grid is None else self.grid.ctypes.data_as(C.c_void_p)
        return (self.data, self.nmb, self.n, self.lo, self.hi, self.v0, self.dx, gp, self.G,
                self.g0.ctypes.data_as(C.c_void_p), self.gh)

    def sample_scalars(self, S, a, fallback=np.pi / 3, nthreads=None, with_blocks=False):
        S = np.ascontiguousarray(S, dtype=np.float64)
        ns = S.size // 8
        out = np.empty((ns, 5))
        mb = np.empty(ns, dtype=np.int64) if with_blocks else None
        lib().mko_sample_scalars(*self._args(), S.reshape(-1), ns, a, fallback, out.reshape(-1),
                                 None if mb is None else mb.ctypes.data_as(C.c_void_p), threads(nthreads))
        out = out.reshape(S.shape[:-1] + (5,))
        return (out, mb.reshape(S.shape[:-1])) if with_blocks else out

    def render(self, a, inc, dist, fov, res, max_nsteps, fluid_gamma, r_high, mass_scale, M_bh, nu_obs,
               p0=0, count=None, stride=1, nthreads=None, force_steps=None, margins=False):
        """Pixels p0 + q * stride (q < count) of the res x res image -> (I, steps).

        force_steps: accepted steps per ray to run instead of the freeze rule
        (a candidate the rule rejects is accepted with the unclamped step).
        margins: also return (count, 2) relative distances of the rejected
        candidate and of the last accepted state from the nearest freeze
        threshold (|r - r_EH| against tol or 1500; inf where not defined).
        """
        count = (res * res - p0 + stride - 1) // stride if count is None else count
        I = np.empty(count)
        steps = np.empty(count, dtype=np.int32)
        grid = self.grid if self.grid is not None else np.full((1, 1, 1), -1, dtype=np.int32)
        fs = None
        if force_steps is not None:
            fs = np.ascontiguousarray(force_steps, dtype=np.int32).reshape(-1)
            assert fs.size == count
        mg = np.empty((count, 2)) if margins else None
        lib().mko_render(self.data, self.nmb, self.n, self.lo, self.hi, self.v0, self.dx, grid, self.G,
                         self.g0, self.gh, a, inc, dist, fov, res, max_nsteps, fluid_gamma, r_high, mass_scale,
                         M_bh, nu_obs, p0, count, stride, I, steps,
                         None if fs is None else fs.ctypes.data_as(C.c_void_p),
                         None if mg is None else mg.ctypes.data_as(C.c_void_p), threads(nthreads))
        return (I, steps, mg) if margins else (I, steps)


def synchrotron(Ne, Th, B, pitch, nu, invariant=True, rescale=1.0):
    arrs = [np.ascontiguousarray(v, dtype=np.float64) for v in np.broadcast_arrays(Ne, Th, B, pitch, nu)]
    j = np.empty(arrs[0].shape)
    al = np.empty(arrs[0].shape)
    lib().mko_synchrotron(*[a.reshape(-1) for a in arrs], arrs[0].size, int(bool(invariant)), rescale,
                          j.reshape(-1), al.reshape(-1))
    return j, al
