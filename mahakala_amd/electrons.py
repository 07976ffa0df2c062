"""Electron temperature model (reference: mahakala/electrons.py:23-50)."""
from __future__ import annotations

import numpy as np

from mahakala_amd import _lib
from mahakala_amd._array import wrap

KB = 1.3807e-16
CL = 2.99792458e10
ME = 9.1094e-28
MP = 1.6726e-24
EC = 4.8032e-10
HPL = 6.6261e-27
GNEWT = 6.6743e-8


def rlow_rhigh_model(dens, u, beta, r_low=1, r_high=40, electron_gamma=4. / 3, ion_gamma=5. / 3):
    """Dimensionless electron temperature from the R(beta) model.

    T_ratio = (r_high beta^2 + r_low) / (1 + beta^2) and
    Theta_e = c^2 m_p u (g_e - 1)(g_i - 1) / (rho ((g_i - 1) + (g_e - 1) T_ratio)) / (m_e c^2),
    elementwise on the GPU (electrons.py:32-50).
    """
    arrs = np.broadcast_arrays(*[np.asarray(v, dtype=np.float64) for v in (dens, u, beta)])
    shape = arrs[0].shape
    n = int(np.prod(shape)) if shape else 1
    bufs = [_lib.DeviceBuffer.from_array(a) for a in arrs]
    out = _lib.DeviceBuffer(8 * n)
    _lib.check(_lib.lib().mk_rlow_rhigh(bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, n, float(r_low), float(r_high),
                                        float(electron_gamma), float(ion_gamma), out.ptr, None))
    return wrap(out.to_array(shape))
