"""Pins of the oracle's image glue and polar camera (CPU only).

Reference: images.py:86-118 (beta, sigma, Theta_e, units, local frequency,
invariant coefficients, sigma cut, transfer) and geodesics.py:107-134.

The glue is pinned at a point where the fluid quantities have closed forms:
a Schwarzschild hole (a = 0), a fluid at rest (AthenaK velocity 0, so u is
the normal observer), a field along z, sampled in the equatorial plane with
k = (1, 1, 0, 0).  There g^{tt} = -(1 + f) with f = 2/r, so the lapse is
alpha = (1 + 2/r)^(-1/2), k.u = -alpha, b = alpha B_z and k.b = 0 (pitch
angle pi/2): the local frequency is alpha nu_obs, the plasma beta is
(gamma - 1) u / (b^2 / 2), and one sample contributes -dt L j.
"""
import numpy as np
import pytest

from oracle import mahakala_np as onp

def _uniform_block(center, width, n, dens, u, bz):
    x1f = center[0] + np.linspace(-width / 2, width / 2, n + 1)
    x2f = center[1] + np.linspace(-width / 2, width / 2, n + 1)
    x3f = center[2] + np.linspace(-width / 2, width / 2, n + 1)
    v = [0.5 * (f[1:] + f[:-1]) for f in (x1f, x2f, x3f)]
    uov = np.zeros((5, 1, n, n, n))
    uov[0] = dens
    uov[4] = u
    B = np.zeros((3, 1, n, n, n))
    B[2] = bz
    blocks = onp.build_meshblocks(uov, B, np.array([0]), np.array([[0, 0, 0]]))
    names = ["dens", "velx", "vely", "velz", "eint"]
    return onp.Snapshot(blocks, v[0][None], v[1][None], v[2][None], x1f[None], x2f[None], x3f[None],
                        names + ["bcc1", "bcc2", "bcc3"], 0.0, 13. / 9)


def _history(x0, rows, dt):
    """Rows at (x0, 0, 0) with k = (1, 1, 0, 0); dt per row (the glue does not
    need the rows to move)."""
    S = np.zeros((rows, 1, 8))
    S[:, 0, 1] = x0
    S[:, 0, 4] = 1.0
    S[:, 0, 5] = 1.0
    return S, np.full((rows, 1), dt)


@pytest.mark.parametrize("nu_obs", [86e9, 230e9])
def test_image_glue_static_fluid_closed_form(nu_obs):
    r0, dens, u, bz, gam = 1.0e4, 1.3, 0.4, 0.9, 13. / 9
    snap = _uniform_block((r0, 0.0, 0.0), 8.0, 4, dens, u, bz)
    alpha = (1.0 + 2.0 / r0) ** -0.5
    b = alpha * bz
    beta = (gam - 1.0) * u / (0.5 * b * b)
    units = onp.get_units(6.2e9 * onp.Msun, 1e26)
    theta = onp.rlow_rhigh_model(dens, u, beta, r_high=40)
    j, al = onp.synchrotron_coefficients(units["Ne_unit"] * dens, theta, units["B_unit"] * b, np.pi / 2,
                                         alpha * nu_obs, invariant=True, rescale_nu=1.0 / nu_obs)
    assert j > 0 and al > 0
    # the pieces the glue computes, at the sample
    S, dt = _history(r0, 2, -0.02)
    fs = onp.get_fluid_scalars_from_geodesics(snap, S[1:])
    np.testing.assert_allclose(fs["kdotu"], -alpha, rtol=1e-13)
    np.testing.assert_allclose(fs["b"], b, rtol=1e-13)
    np.testing.assert_allclose(fs["pitch_angle"], np.pi / 2, rtol=1e-13)
    # one contributing row: I = -dt[0] L j[1]
    I = onp.image_from_history(snap, S, dt, gam, observing_frequency=nu_obs)
    np.testing.assert_allclose(I, -dt[0, 0] * units["L_unit"] * j, rtol=1e-11)
    # two contributing rows of equal coefficients: c j (1 + (1 - c alpha))
    S, dt = _history(r0, 3, -0.02)
    I = onp.image_from_history(snap, S, dt, gam, observing_frequency=nu_obs)
    c = 0.02 * units["L_unit"]
    np.testing.assert_allclose(I, c * j * (2.0 - c * al), rtol=1e-11)


def test_image_glue_sigma_cut_and_empty_cells():
    r0, gam = 1.0e4, 13. / 9
    S, dt = _history(r0, 2, -0.02)
    # b^2 / rho just above 100: j and alpha are set to 0
    bz = np.sqrt(100.5 * 1.3) / (1.0 + 2.0 / r0) ** -0.5
    snap = _uniform_block((r0, 0.0, 0.0), 8.0, 4, 1.3, 0.4, bz)
    assert onp.image_from_history(snap, S, dt, gam)[0] == 0.0
    # just below the cut the sample contributes
    snap = _uniform_block((r0, 0.0, 0.0), 8.0, 4, 1.3, 0.4, 0.99 * bz)
    assert onp.image_from_history(snap, S, dt, gam)[0] > 0.0
    # zero density (and no field): NaN coefficients become 0
    snap = _uniform_block((r0, 0.0, 0.0), 8.0, 4, 0.0, 0.0, 0.0)
    assert onp.image_from_history(snap, S, dt, gam)[0] == 0.0
    # outside every block: zero primitives, no contribution
    S_out, _ = _history(-r0, 2, -0.02)
    assert onp.image_from_history(snap, S_out, dt, gam)[0] == 0.0


@pytest.mark.parametrize("inc", [0.0, 17.0, 60.0])
def test_polar_camera_pixels_closed_form(inc):
    radius = np.array([0.5, 2.0, 5.0, 9.5])
    angle = np.array([0.0, 1.0, 3.5, 5.9])
    sx, sv = onp.get_camera_pixel(inc, 1000.0, radius, angle)
    i, d = np.deg2rad(inc), 1000.0
    x, y = radius * np.cos(angle), radius * np.sin(angle)
    np.testing.assert_allclose(sx[1], -y * np.cos(i) + d * np.sin(i), rtol=1e-14, atol=1e-12)
    np.testing.assert_allclose(sx[2], x, rtol=1e-14, atol=1e-14)
    np.testing.assert_allclose(sx[3], y * np.sin(i) + d * np.cos(i), rtol=1e-14, atol=1e-12)
    k = np.stack(sv[1:], 1)
    khat = k / np.linalg.norm(k, axis=1)[:, None]
    np.testing.assert_allclose(khat, np.broadcast_to([np.sin(i), 0, np.cos(i)], khat.shape), atol=1e-12)
