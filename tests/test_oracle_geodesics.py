"""Pins of the oracle's camera, integrator, freeze/trim rules and photon
selection (CPU only).  Reference: geodesics.py:29-378."""
import os

import numpy as np
import pytest

from oracle import c_oracle
from oracle import mahakala_np as onp

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.mark.parametrize("a,inc", [(0.9, 60.0), (0.0, 30.0), (-0.5, 85.0)])
def test_camera_closed_form(a, inc):
    n, d = 6, 1000.0
    s0 = onp.initialize_geodesics_at_camera(a, inc, d, -10, 10, n)
    grid = -10 + (2 * np.arange(n) + 1) * (20.0 / (2 * n))
    X, Y = np.meshgrid(grid, grid, indexing="ij")
    i = np.deg2rad(inc)
    pos = np.stack([-Y.ravel() * np.cos(i) + d * np.sin(i), X.ravel(), Y.ravel() * np.sin(i) + d * np.cos(i)], 1)
    np.testing.assert_allclose(s0[:, 1:4], pos, rtol=1e-14, atol=1e-12)
    assert np.all(s0[:, 0] == 0) and np.all(s0[:, 4] == 1)
    # spatial k is parallel to (sin i, 0, cos i), pointing away from the hole
    k = s0[:, 5:]
    khat = k / np.linalg.norm(k, axis=1)[:, None]
    np.testing.assert_allclose(khat, np.broadcast_to([np.sin(i), 0, np.cos(i)], khat.shape), atol=1e-12)
    # null
    g = onp.metric(s0[:, :4], a)
    kk = np.einsum("ni,nij,nj->n", s0[:, 4:], g, s0[:, 4:])
    np.testing.assert_allclose(kk, 0.0, atol=1e-13)
    np.testing.assert_allclose(c_oracle.camera_grid(a, inc, d, -10, 10, n), s0, rtol=1e-14, atol=1e-12)


def test_odd_grid_has_nan_centre_pixel():
    s0 = onp.initialize_geodesics_at_camera(0.5, 30, 1000, -10, 10, 5)
    assert np.isnan(s0[12, 5:]).all() and np.isfinite(np.delete(s0, 12, 0)).all()


def _invariants(S, a):
    x, v = S[..., :4], S[..., 4:]
    g = onp.metric(x.reshape(-1, 4), a).reshape(x.shape[:-1] + (4, 4))
    p = np.einsum("...ij,...j->...i", g, v)
    H = 0.5 * np.einsum("...i,...i->...", p, v)
    E = -p[..., 0]
    Lz = x[..., 1] * p[..., 2] - x[..., 2] * p[..., 1]
    r = onp.radius_cal(x, a)
    c = x[..., 3] / r
    s = np.sqrt(1 - c * c)
    pth = (x[..., 1] * p[..., 1] + x[..., 2] * p[..., 2]) * c / s - r * s * p[..., 3]
    Q = pth ** 2 + c * c * (Lz ** 2 / (s * s) - a * a * E * E)
    return H, E, Lz, Q


def test_conserved_quantities_along_rays():
    """H = 0, E = -p_t, L_z = p_phi and the Carter constant Q stay constant
    along the integrated rays (RK4 truncation only)."""
    a = 0.9
    s0 = c_oracle.camera_grid(a, 60, 1000, -10, 10, 6)
    S, D = c_oracle.geodesic_integrator(10000, s0, 40, 1e-4, a)
    H, E, Lz, Q = _invariants(S, a)
    live = (D != 0) & (onp.radius_cal(S, a) > 2.5)
    assert np.abs(H[live]).max() < 1e-5
    assert np.abs((E - E[0])[live]).max() < 1e-6
    assert np.abs((Lz - Lz[0])[live]).max() < 1e-6
    assert np.abs((Q - Q[0])[live]).max() < 1e-5 * np.abs(Q[0]).max()


def test_rk4_is_fourth_order():
    a = 0.9
    s0 = c_oracle.camera_grid(a, 60, 1000, -10, 10, 6)
    S, _ = c_oracle.geodesic_integrator(10000, s0, 40, 1e-4, a)
    r = onp.radius_cal(S, a)
    st = S[np.argmin(np.abs(r[:, 20] - 8)), 20:21]

    def run(n):
        s = st.copy()
        for _ in range(n):
            s = c_oracle.rk4(s, -3.0 / n, a)
        return s
    ref = run(3200)
    e = [np.abs(run(n) - ref).max() for n in (10, 20, 40)]
    assert 14 < e[0] / e[1] < 18 and 14 < e[1] / e[2] < 18
    # the NumPy (jacfwd + LU) and C (closed form) RK4 agree on one step
    np.testing.assert_allclose(onp.RK4_gen(st, np.array([-0.3]), a), c_oracle.rk4(st, -0.3, a), rtol=1e-13)


def _reference_trim(dt, N):
    """The trim of geodesics.py:275-281 applied to a full (N, npx) dt array."""
    first = int(np.argmax(np.all(dt == 0, axis=1)))
    if first < 1:
        first = N
    return dt[:first + 2].shape[0]


def test_trim_rule_brute_force():
    rng = np.random.Generator(np.random.PCG64(0))
    for _ in range(200):
        N = int(rng.integers(2, 30))
        npx = int(rng.integers(1, 6))
        K = rng.integers(0, N + 1, npx)           # K = N: never froze
        dt = np.where(np.arange(N)[:, None] < K[None, :], -1.0, 0.0)
        assert c_oracle.trim_rows(K, N) == _reference_trim(dt, N)


def test_numpy_and_c_integrators_agree():
    f = np.load(os.path.join(GOLD, "geodesic_6x6.npz"))
    a = float(f["spin"])
    K, last, _ = c_oracle.integrate_final(f["s0"], 10000, 40, 1e-4, a)
    np.testing.assert_array_equal(K, f["K"])
    np.testing.assert_allclose(last, f["S_last"], rtol=1e-8, atol=1e-8)
    S, D = c_oracle.geodesic_integrator(10000, f["s0"], 40, 1e-4, a)
    assert S.shape[0] == int(f["T"])
    np.testing.assert_allclose(S[f["rows"]], f["S_rows"], rtol=1e-8, atol=1e-8)
    # frozen rows repeat the last accepted state with dt = 0
    for p in range(S.shape[1]):
        k = int(f["K"][p])
        assert np.all(D[k:, p] == 0) and np.all(D[:k, p] != 0)
        assert np.all(S[k:, p] == S[k, p])


def test_select_photons_reads_row_before_argmax():
    """select_photons_integrator's value is r one row before argmax(dt); the
    two .at lines of the reference are no-ops (geodesics.py:373-378)."""
    a, inc = 0.9, 60.0
    ang = np.linspace(0.1, 3.0, 6)
    rad = np.full(6, 5.0)
    sx, sv = onp.get_camera_pixel(inc, 1000, rad, ang)
    S, dt = onp.geodesic_integrator(2000, onp.initial_condition(sx, sv, a), 40, 1e-2, a)
    r = onp.radius_cal(S, a)
    m = np.argmax(dt, axis=0)
    expect = r[m - 1, np.arange(6)]
    np.testing.assert_array_equal(onp.select_photons_integrator(inc, ang, rad, a), expect)
    _, _, sel = c_oracle.integrate_final(onp.initial_condition(sx, sv, a), 2000, 40, 1e-2, a)
    np.testing.assert_allclose(sel, expect, rtol=1e-8)


@pytest.mark.parametrize("a", [0.0, 0.7])
def test_equator_camera_closed_form(a):
    """The 'equator' camera (geodesics.py:183-197): a line of rays at x = d,
    y = the pixel centres linspace(lo, hi, 2n+1)[1::2] = lo + (2k+1)(hi-lo)/2n,
    z = 0 whatever the inclination, all with k = (1, 1, 0, 0) before
    normalisation; after nullify (geodesics.py:58-85) k^t = 1, k^y = k^z = 0,
    k^x > 0 and g(k, k) = 0."""
    n, lo, hi, d = 60, -15.0, 15.0, 1000.0
    for inc in (0.0, 60.0):
        x, v = onp.get_initial_grid(inc, d, lo, hi, n, "equator")
        yk = lo + (2 * np.arange(n) + 1) * (hi - lo) / (2 * n)
        np.testing.assert_array_equal(x[0], 0.0)
        np.testing.assert_array_equal(x[1], d)
        np.testing.assert_allclose(x[2], yk, rtol=0, atol=1e-13)
        np.testing.assert_array_equal(x[3], 0.0)
        np.testing.assert_array_equal(v, np.array([[1.0] * n, [1.0] * n, [0.0] * n, [0.0] * n]))
    s = onp.initialize_geodesics_at_camera(a, 60, d, lo, hi, n, camera_type="equator")
    assert s.shape == (n, 8)
    np.testing.assert_array_equal(s[:, 4], 1.0)
    np.testing.assert_array_equal(s[:, 6:], 0.0)
    assert np.all(s[:, 5] > 0)
    for q in range(n):
        g = onp.metric(s[q, :4], a)
        k = s[q, 4:]
        assert abs(k @ g @ k) < 1e-12
