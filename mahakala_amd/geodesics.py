"""Null geodesics in Cartesian Kerr-Schild spacetime on the GPU.

Drop-in for the reference's ``mahakala/geodesics.py``: same names, arguments,
defaults and return shapes.  Host-side camera geometry (``get_initial_grid``,
``_Image_to_BH``, ``_perpendicular``, ``get_camera_pixel``) is host NumPy, as
it is in the reference; every numerical kernel (nullify, metric, inverse,
geodesic RHS, RK4, radius, integrator, photon selection, bisection) runs in
libmahakala_hip.so.  Arrays come back as ``MkArray`` (NumPy with ``.at``).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from mahakala_amd import _lib
from mahakala_amd._array import wrap

__all__ = [
    "initialize_geodesics_at_camera", "metric", "imetric", "get_camera_pixel", "get_initial_grid",
    "initial_condition", "geodesic_integrator", "radius_cal", "rhs", "RK4_gen", "radius_EH",
    "select_photons_integrator", "find_shadow_bisection", "find_shadow_bisection_angles",
]


def _dev(a):
    return _lib.DeviceBuffer.from_array(a)


def initialize_geodesics_at_camera(bhspin, inclination, distance, fov_lower, fov_upper, pixels_per_side,
                                   camera_type='grid'):
    """Null camera rays (npx, 8) = [t, x, y, z, k^t, k^x, k^y, k^z]  (geodesics.py:29-55).

    'grid' builds the pixel grid and null-normalises it in one kernel;
    'equator' builds the line of rays on the host and normalises on the GPU.
    """
    if isinstance(camera_type, str) and camera_type.lower() == 'grid':
        n = int(pixels_per_side)
        out = _lib.DeviceBuffer(8 * 8 * n * n)
        _lib.check(_lib.lib().mk_camera_grid(float(bhspin), float(inclination), float(distance), float(fov_lower),
                                             float(fov_upper), n, out.ptr, None))
        return wrap(out.to_array((n * n, 8)))
    s0_x, s0_v = get_initial_grid(inclination, distance, fov_lower, fov_upper, pixels_per_side, camera_type)
    return initial_condition(s0_x, s0_v, bhspin)


def _Nullify(metric_fn, bhspin, p=1):
    """Per-point nullify factory (geodesics.py:69-85); k^t is held fixed."""
    assert p > 0
    if p != 1:
        raise NotImplementedError("only p=1 (k^t fixed) is used by the reference")

    def nullify(x, v):
        out = initial_condition(np.asarray(x, dtype=np.float64).reshape(4, 1),
                                np.asarray(v, dtype=np.float64).reshape(4, 1), bhspin)
        return wrap(np.asarray(out)[0, 4:])

    return nullify


def metric(x, bhspin):
    """Covariant KS metric at x (..., 4) -> (..., 4, 4)  (geodesics.py:88-104)."""
    x = np.asarray(x, dtype=np.float64)
    lead = x.shape[:-1]
    n = int(np.prod(lead)) if lead else 1
    d_x, out = _dev(x), _lib.DeviceBuffer(8 * 16 * n)
    _lib.check(_lib.lib().mk_metric(d_x.ptr, n, float(bhspin), out.ptr, None))
    return wrap(out.to_array(lead + (4, 4)))


def imetric(x, bhspin):
    """Contravariant metric (..., 4, 4), closed form eta - f l l  (geodesics.py:339-347)."""
    x = np.asarray(x, dtype=np.float64)
    lead = x.shape[:-1]
    n = int(np.prod(lead)) if lead else 1
    d_x, out = _dev(x), _lib.DeviceBuffer(8 * 16 * n)
    _lib.check(_lib.lib().mk_imetric(d_x.ptr, n, float(bhspin), out.ptr, None))
    return wrap(out.to_array(lead + (4, 4)))


def _Image_to_BH(x, y, z, i, d):
    """Image-plane point -> BH frame  (geodesics.py:204-209)."""
    i = i * np.pi / 180
    x_BH = -y * np.cos(i) + z * np.sin(i) + d * np.sin(i)
    y_BH = x
    z_BH = y * np.sin(i) + z * np.cos(i) + d * np.cos(i)
    return np.array([x_BH, y_BH, z_BH])


def _perpendicular(a):
    """(x, y) -> (x + y, y - x)  (geodesics.py:212-216)."""
    b = np.zeros_like(np.asarray(a, dtype=np.float64))
    b[0] = a[0] + a[1]
    b[1] = a[1] - a[0]
    return b


def _plane_rays(x, y, inclination, distance):
    n = np.size(x)
    origin = _Image_to_BH(0, 0, 0, inclination, distance)
    perp = _perpendicular([x, y])
    init = _Image_to_BH(x, y, np.zeros(n), inclination, distance)
    perp_bh = _Image_to_BH(perp[0], perp[1], np.zeros(n), inclination, distance)
    k = np.cross(-init.T + origin, (perp_bh - init).T)
    s0_x = np.array([np.zeros(n), init[0].ravel(), init[1].ravel(), init[2].ravel()])
    s0_v = np.array([np.ones(n), k.T[0].ravel(), k.T[1].ravel(), k.T[2].ravel()])
    return s0_x, s0_v


def get_camera_pixel(inclination, distance, radius, angle):
    """Positions and (un-normalised) wavevectors at polar image points  (geodesics.py:107-134)."""
    size = np.size(radius)
    x = np.ones(size) * np.cos(angle) * radius
    y = np.ones(size) * np.sin(angle) * radius
    return _plane_rays(x, y, inclination, distance)


def get_initial_grid(inclination, distance, fov_lower, fov_upper, spacing, camera_type):
    """Camera grid ('grid') or equatorial line ('equator')  (geodesics.py:137-201)."""
    kind = camera_type.lower()
    if kind == 'grid':
        grid = np.linspace(fov_lower, fov_upper, 2 * spacing + 1)[1::2]
        x, y = np.meshgrid(grid, grid, indexing='ij')
        return _plane_rays(x.ravel(), y.ravel(), inclination, distance)
    if kind == 'equator':
        grid = np.linspace(fov_lower, fov_upper, 2 * spacing + 1)[1::2]
        s0_x = np.zeros((4, len(grid)))
        s0_x[1] = distance
        s0_x[2] = grid
        s0_v = np.ones((4, len(grid)))
        s0_v[2] = 0
        s0_v[3] = 0
        return s0_x, s0_v
    print(f'Unexpected camera type "{camera_type}".')
    print('Please choose either "grid" or "equator"')
    return None


def initial_condition(s0_x, s0_v, bhspin):
    """Make wavevectors null at fixed k^t: (4, n), (4, n) -> (n, 8)  (geodesics.py:219-230)."""
    x = np.ascontiguousarray(np.asarray(s0_x, dtype=np.float64).T)
    v = np.ascontiguousarray(np.asarray(s0_v, dtype=np.float64).T)
    n = x.shape[0]
    d_x, d_v, out = _dev(x), _dev(v), _lib.DeviceBuffer(8 * 8 * n)
    _lib.check(_lib.lib().mk_initial_condition(d_x.ptr, d_v.ptr, n, float(bhspin), out.ptr, None))
    return wrap(out.to_array((n, 8)))


# Bytes of one materialised history row of one ray: state (8) and dt, FP64.
HISTORY_ROW_BYTES = 9 * 8
# Share of the free device memory a history chunk may take when no explicit
# budget is given.
DEVICE_BUDGET_FRACTION = 0.5


def history_chunk_rays(n, T, max_chunk_bytes=None):
    """Rays per chunk of a materialised (T, n) history: as many as fit
    max_chunk_bytes of device memory (T rows of 72 bytes per ray), or, when it
    is None, half of the free device memory; at least one."""
    budget = max_chunk_bytes
    if budget is None:
        free, _ = _lib.device_memory()
        budget = int(free * DEVICE_BUDGET_FRACTION)
    per_ray = max(1, int(T)) * HISTORY_ROW_BYTES
    return int(max(1, min(n, budget // per_ray)))


def geodesic_integrator(N, s0, div, tol, bhspin, max_chunk_bytes=None):
    """RK4 geodesics with the radius-scaled step  (geodesics.py:233-281).

    Returns (S (T, npx, 8), final_dt (T, npx)) with the reference's fixed-N
    scan semantics: a ray is frozen at its first rejected step and repeats
    (state, 0) afterwards; T = argmax(all frozen) + 2, or N if that is < 1.
    The GPU integrates each ray only until it freezes: one pass over all rays
    finds T, then the T rows are written in chunks of rays whose history fits
    max_chunk_bytes of device memory (None: half of the free device memory),
    so a history larger than the GPU's memory is still returned whole.
    Chunking changes no value: every ray is integrated alone.
    """
    s0 = np.ascontiguousarray(np.asarray(s0, dtype=np.float64))
    n = s0.shape[0]
    N = int(N)
    d_s0 = _dev(s0)
    d_K = _lib.DeviceBuffer(4 * n)
    L = _lib.lib()
    _lib.check(L.mk_integrate_final(d_s0.ptr, n, N, float(div), float(tol), float(bhspin), d_K.ptr, None, None,
                                    None))
    T = C.c_long()
    _lib.check(L.mk_trim_rows(d_K.ptr, n, N, C.byref(T), None))
    T = T.value
    del d_K
    m = history_chunk_rays(n, T, max_chunk_bytes) if n else 1
    S = np.empty((T, n, 8))
    D = np.empty((T, n))
    d_S = _lib.DeviceBuffer(8 * 8 * T * m)
    d_D = _lib.DeviceBuffer(8 * T * m)
    for lo in range(0, n, m):
        k = min(m, n - lo)
        _lib.check(L.mk_integrate_history(d_s0.ptr + 8 * 8 * lo, k, N, float(div), float(tol), float(bhspin), T,
                                          d_S.ptr, d_D.ptr, None))
        if k == n:
            _lib.check(L.mk_memcpy_d2h(S.ctypes.data, d_S.ptr, S.nbytes, None))
            _lib.check(L.mk_memcpy_d2h(D.ctypes.data, d_D.ptr, D.nbytes, None))
        else:
            S[:, lo:lo + k] = d_S.to_array((T, k, 8))
            D[:, lo:lo + k] = d_D.to_array((T, k))
    return wrap(S), wrap(D)


def radius_cal(x, bhspin):
    """Spherical KS radius of Cartesian KS points (..., >=4)  (geodesics.py:284-291)."""
    x = np.asarray(x, dtype=np.float64)
    lead = x.shape[:-1]
    stride = x.shape[-1]
    n = int(np.prod(lead)) if lead else 1
    d_x, out = _dev(x), _lib.DeviceBuffer(8 * n)
    _lib.check(_lib.lib().mk_radius_cal(d_x.ptr, n, stride, float(bhspin), out.ptr, None))
    r = out.to_array(lead)
    return wrap(r) if lead else float(r)


def rhs(state1, bhspin):
    """[v, -Gamma v v] for one (8,) or many (..., 8) states  (geodesics.py:294-309)."""
    s = np.asarray(state1, dtype=np.float64)
    lead = s.shape[:-1]
    n = int(np.prod(lead)) if lead else 1
    d_s, out = _dev(s), _lib.DeviceBuffer(8 * 8 * n)
    _lib.check(_lib.lib().mk_rhs(d_s.ptr, n, float(bhspin), out.ptr, None))
    return wrap(out.to_array(lead + (8,)))


def _vectorized_rhs(s0, bhspin):
    """rhs over rays  (geodesics.py:312-314)."""
    return rhs(s0, bhspin)


def RK4_gen(state1, dt, bhspin):
    """One classic RK4 step per ray: (n, 8), (n,) -> (n, 8)  (geodesics.py:317-336)."""
    s = np.ascontiguousarray(np.asarray(state1, dtype=np.float64))
    n = s.shape[0]
    dt = np.ascontiguousarray(np.broadcast_to(np.asarray(dt, dtype=np.float64).reshape(-1), (n,)))
    d_s, d_dt, out = _dev(s), _dev(dt), _lib.DeviceBuffer(8 * 8 * n)
    _lib.check(_lib.lib().mk_rk4(d_s.ptr, d_dt.ptr, n, float(bhspin), out.ptr, None))
    return wrap(out.to_array((n, 8)))


def radius_EH(a_spin):
    """Outer horizon radius 1 + sqrt(1 - a^2)  (geodesics.py:350-351)."""
    return 1 + np.sqrt(1 - a_spin ** 2)


def select_photons_integrator(inc, angle, radius, bhspin, distance=1000, max_steps=2000):
    """Final radius of each (radius, angle) camera ray  (geodesics.py:354-378).

    As in the reference, the value is r one row before argmax(dt) of the
    ray's history (its last accepted state's predecessor), with index -1
    wrapping to the last row.
    """
    radius = np.ascontiguousarray(np.asarray(radius, dtype=np.float64).reshape(-1))
    n = radius.size
    angle = np.ascontiguousarray(np.broadcast_to(np.asarray(angle, dtype=np.float64).reshape(-1), (n,)))
    out = np.empty(n)
    _lib.check(_lib.lib().mk_select_photons(float(inc), angle.ctypes.data, radius.ctypes.data, n, float(bhspin),
                                            float(distance), int(max_steps), out.ctypes.data))
    return wrap(out)


def find_shadow_bisection(bhspin, inc, num_angles, max_steps=2000, error_allowed=0.001, max_it=40):
    """Shadow boundary at num_angles uniform angles, curve closed  (geodesics.py:381-402)."""
    angles = np.arange(num_angles) / num_angles * 2. * np.pi
    radii = find_shadow_bisection_angles(bhspin, inc, angles, max_it=max_it, error_allowed=error_allowed,
                                         max_steps=max_steps)
    radii = np.append(radii, radii[0])
    angles = np.append(angles, angles[0])
    return angles, radii


def find_shadow_bisection_angles(bhspin, inc, angles, max_steps=2000, error_allowed=0.001, max_it=40):
    """Bisection on the final radius (< 100 falls in) per angle  (geodesics.py:405-435).

    Every angle starts from the bracket [0.5, 10] and is halved once per
    iteration, so all angles run the same number of iterations; the GPU runs
    each angle's whole bisection in its own lane.
    """
    angles = np.ascontiguousarray(np.asarray(angles, dtype=np.float64).reshape(-1))
    out = np.empty(angles.size)
    _lib.check(_lib.lib().mk_shadow_bisect(float(bhspin), float(inc), angles.ctypes.data, angles.size,
                                           int(max_steps), float(error_allowed), int(max_it), 1000.0,
                                           out.ctypes.data))
    return out
