"""Thermal-synchrotron coefficients and radiative transfer on the GPU.

Drop-in for the reference's ``mahakala/transfer.py`` (same names, arguments
and defaults).  All arithmetic is FP64 in libmahakala_hip.so.
"""
from __future__ import annotations

import numpy as np

from mahakala_amd import _lib
from mahakala_amd._array import wrap

__all__ = ["synchrotron_coefficients", "solve_specific_intensity", "solve_attenuated_emissivity"]


def synchrotron_coefficients(Ne, Theta_e, B, pitch_angle, nu, invariant=True, rescale_nu=1.):
    """Thermal synchrotron emissivity and absorptivity  (transfer.py:30-86).

    nu_s = (2/9) (e B / 2 pi m_e c) Theta_e^2 sin(theta), X = nu / nu_s,
    j = Ne nu_s (sqrt X + 2^(11/12) X^(1/6))^2 / (2 Theta_e^2) e^(-X^(1/3)) sqrt2 pi e^2 / 3c,
    zero for X > 1e12 or Theta_e < 0.3; alpha = j / B_nu (Planck, with the
    series form of e^x - 1 below x = 2e-3); invariant forms j / (nu s)^2 and
    alpha nu s with s = rescale_nu; NaN -> 0.  Inputs broadcast together.
    """
    arrs = np.broadcast_arrays(*[np.asarray(v, dtype=np.float64) for v in (Ne, Theta_e, B, pitch_angle, nu)])
    shape = arrs[0].shape
    n = int(np.prod(shape)) if shape else 1
    bufs = [_lib.DeviceBuffer.from_array(a) for a in arrs]
    j = _lib.DeviceBuffer(8 * n)
    al = _lib.DeviceBuffer(8 * n)
    _lib.check(_lib.lib().mk_synchrotron(*[b.ptr for b in bufs], n, int(bool(invariant)), float(rescale_nu), j.ptr,
                                         al.ptr, None))
    return wrap(j.to_array(shape)), wrap(al.to_array(shape))


def _rt_inputs(emissivity, absorptivity, dt):
    j = np.ascontiguousarray(np.asarray(emissivity, dtype=np.float64))
    if j.ndim != 2:
        raise ValueError("emissivity must have shape (nsteps, npx)")
    T, npx = j.shape
    al = np.ascontiguousarray(np.broadcast_to(np.asarray(absorptivity, dtype=np.float64), (T, npx)))
    d = np.ascontiguousarray(np.broadcast_to(np.asarray(dt, dtype=np.float64), (T, npx)))
    return T, npx, [_lib.DeviceBuffer.from_array(a) for a in (j, al, d)]


def solve_specific_intensity(emissivity, absorptivity, dt, L_unit, dIs=False):
    """Explicit-Euler transfer scanned from the far end to the camera  (transfer.py:89-119).

    I <- I - dt[i-1] L (j[i] - alpha[i] I) for i = nsteps-1 .. 1.  With
    ``dIs`` returns (I, dI) where dI (nsteps-1, npx) is in scan order.
    """
    T, npx, (bj, ba, bd) = _rt_inputs(emissivity, absorptivity, dt)
    I = _lib.DeviceBuffer(8 * npx)
    dI = _lib.DeviceBuffer(8 * max(T - 1, 0) * npx) if dIs else None
    _lib.check(_lib.lib().mk_solve_intensity(bj.ptr, ba.ptr, bd.ptr, T, npx, float(L_unit), I.ptr,
                                             dI.ptr if dI is not None else None, None))
    Iv = wrap(I.to_array((npx,)))
    if dIs:
        return Iv, wrap(dI.to_array((max(T - 1, 0), npx)))
    return Iv


def solve_attenuated_emissivity(emissivity, absorptivity, dt, L_unit):
    """Per-step emission attenuated by the optical depth in front of it  (transfer.py:122-144).

    out[i-1] = e^{-tau} (-j[i] dt[i-1] L), then tau -= alpha[i] dt[i-1] L,
    for i = 1 .. nsteps-1 -> (nsteps-1, npx).
    """
    T, npx, (bj, ba, bd) = _rt_inputs(emissivity, absorptivity, dt)
    rows = max(T - 1, 0)
    out = _lib.DeviceBuffer(8 * rows * npx)
    _lib.check(_lib.lib().mk_attenuated_emissivity(bj.ptr, ba.ptr, bd.ptr, T, npx, float(L_unit), out.ptr, None))
    return wrap(out.to_array((rows, npx)))
