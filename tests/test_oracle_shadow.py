"""Pins of the shadow bisection (CPU only).  Reference: geodesics.py:381-435
and tests/test_shadows.py:28-45 (rtol 1e-2 against the Bardeen curve)."""
import os

import numpy as np

from oracle import c_oracle
from oracle import mahakala_np as onp

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_bardeen_fixture_is_the_closed_form():
    f = np.load(os.path.join(GOLD, "shadow_bardeen.npz"))
    for c in range(4):
        a, inc = float(f[f"case{c}_spin"]), float(f[f"case{c}_inc"])
        ang, rad = onp.bardeen_curve(a, inc, np.linspace(*onp.photon_orbit_radii(a), 200001)[1:-1])
        k = np.searchsorted(-ang, -f[f"case{c}_angles"])  # angles run from pi to 0
        k = np.clip(k, 0, ang.size - 1)
        np.testing.assert_allclose(rad[k], f[f"case{c}_radii"], rtol=1e-6)


def test_reference_shadow_cases():
    """The reference's own test at its tolerance, and at 3e-3."""
    f = np.load(os.path.join(GOLD, "shadow_bardeen.npz"))
    for c in range(4):
        a, inc = float(f[f"case{c}_spin"]), float(f[f"case{c}_inc"])
        radii = c_oracle.find_shadow_bisection_angles(a, inc, f[f"case{c}_angles"])
        target = f[f"case{c}_radii"]
        assert np.allclose(radii, target, rtol=1e-2)
        assert np.max(np.abs(radii - target) / target) < 3e-3


def test_schwarzschild_shadow_is_sqrt27():
    for inc in (0.0, 17.0, 60.0, 90.0):
        r = c_oracle.find_shadow_bisection_angles(0.0, inc, np.linspace(0, 6, 8))
        assert np.all(np.abs(r - np.sqrt(27)) < 9.5 / 2 ** 14)


def test_result_is_dyadic_and_iteration_count():
    assert c_oracle.bisect_iterations(1e-3, 40) == 14
    assert c_oracle.bisect_iterations(1e-3, 5) == 5
    r = c_oracle.find_shadow_bisection_angles(0.6, 40.0, np.linspace(0, 6, 16))
    m = (r - 0.5) / (9.5 / 2 ** 14)
    np.testing.assert_array_equal(m, np.round(m))


def test_numpy_oracle_bisection_equals_c():
    f = np.load(os.path.join(GOLD, "shadow_oracle.npz"))
    r = c_oracle.find_shadow_bisection_angles(float(f["spin"]), float(f["inc"]), f["angles"])
    np.testing.assert_array_equal(r, f["radii"])
