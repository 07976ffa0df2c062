"""mahakala_amd -- Mahakala's per-ray hot path on MI355X (gfx950).

Drop-in for the reference package ``mahakala`` (mahakala/__init__.py:22-45):
the same seven top-level functions and the same submodules (geodesics,
transfer, images, electrons, constants, grmhd.grmhd, grmhd.athenak).  The
numerical work runs in hand-written HIP kernels (libmahakala_hip.so, loaded
with ctypes); JAX is not used.  ``import mahakala`` resolves to this package.
"""
from mahakala_amd.geodesics import find_shadow_bisection
from mahakala_amd.geodesics import find_shadow_bisection_angles
from mahakala_amd.geodesics import geodesic_integrator
from mahakala_amd.geodesics import initialize_geodesics_at_camera

from mahakala_amd.transfer import synchrotron_coefficients
from mahakala_amd.transfer import solve_specific_intensity
from mahakala_amd.transfer import solve_attenuated_emissivity

__all__ = [
    "find_shadow_bisection",
    "find_shadow_bisection_angles",
    "geodesic_integrator",
    "initialize_geodesics_at_camera",
    "synchrotron_coefficients",
    "solve_specific_intensity",
    "solve_attenuated_emissivity",
]
