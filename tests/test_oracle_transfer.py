"""Pins of the oracle's units, electron model, synchrotron coefficients and
transfer (CPU only).  Reference: grmhd/grmhd.py:40-55, electrons.py:32-50,
transfer.py:30-144."""
import os

import numpy as np
import pytest

from oracle import c_oracle
from oracle import mahakala_np as onp

GOLD = os.path.join(os.path.dirname(__file__), "golden")
RNG = np.random.Generator(np.random.PCG64(9))


def test_units_closed_form():
    u = onp.get_units(6.2e9 * onp.Msun, 1e26)
    L = onp.GNEWT * 6.2e9 * onp.Msun / onp.CL ** 2
    assert u["L_unit"] == pytest.approx(L, rel=1e-15)
    assert u["T_unit"] * onp.CL == pytest.approx(L, rel=1e-15)
    assert u["dens_unit"] * L ** 3 == pytest.approx(1e26, rel=1e-15)
    assert u["Ne_unit"] * (onp.MP + onp.ME) == pytest.approx(u["dens_unit"], rel=1e-15)
    assert u["B_unit"] ** 2 == pytest.approx(onp.CL ** 2 * 4 * np.pi * u["dens_unit"], rel=1e-14)


def test_rlow_rhigh_limits():
    dens, u = 2.0, 0.3
    # beta -> 0: T_ratio = r_low; beta -> inf: T_ratio = r_high
    for beta, ratio in ((0.0, 1.0), (1e8, 40.0)):
        th = onp.rlow_rhigh_model(dens, u, beta, r_high=40)
        expect = (onp.MP / onp.ME) * u * (1 / 3) * (2 / 3) / (dens * (2 / 3 + ratio / 3))
        assert th == pytest.approx(expect, rel=1e-12)


def _planck(nu, theta):
    return 2 * onp.HPL * nu ** 3 / onp.CL ** 2 / np.expm1(onp.HPL * nu / (onp.ME * onp.CL ** 2 * theta))


def test_kirchhoff_and_invariant_scaling():
    n = 200
    Ne = 10 ** RNG.uniform(3, 7, n)
    Th = 10 ** RNG.uniform(-0.3, 2, n)
    B = 10 ** RNG.uniform(-1, 2, n)
    pitch = RNG.uniform(0.05, np.pi - 0.05, n)
    nu = 10 ** RNG.uniform(10, 12.5, n)
    j, al = onp.synchrotron_coefficients(Ne, Th, B, pitch, nu, invariant=False)
    live = j > 0
    assert live.sum() > 100
    np.testing.assert_allclose((al * _planck(nu, Th))[live], j[live], rtol=1e-9)
    s = 1 / 230e9
    ji, ai = onp.synchrotron_coefficients(Ne, Th, B, pitch, nu, invariant=True, rescale_nu=s)
    np.testing.assert_allclose(ji, j / (nu * s) ** 2, rtol=1e-14)
    np.testing.assert_allclose(ai, al * (nu * s), rtol=1e-14)
    jc, ac = c_oracle.synchrotron(Ne, Th, B, pitch, nu, True, s)
    np.testing.assert_allclose(jc, ji, rtol=1e-13)
    np.testing.assert_allclose(ac, ai, rtol=1e-13)


def test_cut_branches_and_nan():
    base = dict(Ne=1e5, Theta_e=10.0, B=10.0, pitch_angle=1.0, nu=230e9)
    j, _ = onp.synchrotron_coefficients(**{**base, "Theta_e": 0.29})
    assert j == 0
    j, _ = onp.synchrotron_coefficients(**{**base, "nu": 1e40})
    assert j == 0
    j, a = onp.synchrotron_coefficients(**{**base, "B": np.nan})
    assert j == 0 and a == 0
    j, a = onp.synchrotron_coefficients(**{**base, "Ne": 0.0, "B": 0.0})
    assert j == 0 and a == 0


def test_series_matches_expm1_at_the_switch():
    # x just below and above 2e-3 give the same B_nu to the series' accuracy
    th = 0.31  # strong field so that e^{-X^1/3} does not underflow at these nu
    for x in (1.999e-3, 2.001e-3):
        nu = x * onp.ME * onp.CL ** 2 * th / onp.HPL
        j, al = onp.synchrotron_coefficients(1e10, th, 1e7, np.pi / 2, nu, invariant=False)
        assert j > 0
        assert j / al == pytest.approx(_planck(nu, th), rel=1e-11)


def test_transfer_constant_coefficients_closed_form():
    T, npx, L = 50, 4, 3.0
    j = np.full((T, npx), 0.7)
    al = np.full((T, npx), 0.02)
    dt = np.full((T, npx), -0.1)
    c = 0.1 * L                                   # c = -dt L
    I = onp.solve_specific_intensity(j, al, dt, L)
    m = T - 1
    np.testing.assert_allclose(I, 0.7 * (1 - (1 - c * 0.02) ** m) / 0.02, rtol=1e-12)
    I0 = onp.solve_specific_intensity(j, np.zeros_like(al), dt, L)
    np.testing.assert_allclose(I0, m * c * 0.7, rtol=1e-12)   # Riemann sum
    att = onp.solve_attenuated_emissivity(j, al, dt, L)
    k = np.arange(m)[:, None]
    np.testing.assert_allclose(att, np.exp(-k * c * 0.02) * c * 0.7 * np.ones((1, npx)), rtol=1e-12)


def test_backward_scan_equals_forward_accumulation():
    """I = sum_i c_i j_i prod_{k<i} (1 - c_k alpha_k): the form the fused
    kernel accumulates along the ray."""
    f = np.load(os.path.join(GOLD, "transfer.npz"))
    L = float(f["L"])
    j, al, dt = f["jj"], f["aa"], f["dt"]
    I = onp.solve_specific_intensity(j, al, dt, L)
    c = -dt[:-1] * L
    trans = np.cumprod(np.vstack([np.ones((1, j.shape[1])), 1 - c * al[1:]]), axis=0)[:-1]
    fwd = np.sum(trans * c * j[1:], axis=0)
    np.testing.assert_allclose(fwd, I, rtol=1e-12)


def test_golden_transfer_fixture_regenerates():
    f = np.load(os.path.join(GOLD, "transfer.npz"))
    j, al = onp.synchrotron_coefficients(f["Ne"], f["Th"], f["B"], f["pitch"], f["nu"], True, 1 / 230e9)
    np.testing.assert_array_equal(j, f["j"])
    np.testing.assert_array_equal(al, f["al"])
    I, dI = onp.solve_specific_intensity(f["jj"], f["aa"], f["dt"], float(f["L"]), dIs=True)
    np.testing.assert_array_equal(I, f["I"])
    np.testing.assert_array_equal(dI, f["dI"])


def test_dIs_rows_are_in_scan_order_closed_form():
    """solve_specific_intensity(dIs=True) returns the per-step increments in
    the order of the reference's scan, i = T-1 down to 1 (transfer.py:111-117):
    row k is the increment of sample i = T-1-k.  Without absorption that
    increment is -dt[i-1] L j[i], distinct for every (i, pixel) here, so a
    reversed, shifted or transposed row order fails; the rows sum to I."""
    T, npx, L = 7, 3, 2.0
    i = np.arange(T)[:, None]
    p = np.arange(npx)[None, :]
    j = (i + 1.0) + 0.1 * p                     # j[i, p]
    dt = -(i + 1.0) * 0.01 - 0.001 * p          # dt[i, p]
    I, dI = onp.solve_specific_intensity(j, np.zeros((T, npx)), dt, L, dIs=True)
    assert dI.shape == (T - 1, npx)
    for k in range(T - 1):
        ii = T - 1 - k
        np.testing.assert_array_equal(dI[k], -dt[ii - 1] * L * j[ii])
    np.testing.assert_allclose(dI.sum(0), I, rtol=1e-15)
    # with absorption, row k uses the intensity accumulated over rows < k
    al = 0.3 + 0.01 * i + 0 * p
    I2, dI2 = onp.solve_specific_intensity(j, al, dt, L, dIs=True)
    acc = np.zeros(npx)
    for k in range(T - 1):
        ii = T - 1 - k
        np.testing.assert_allclose(dI2[k], -dt[ii - 1] * L * (j[ii] - al[ii] * acc), rtol=1e-14)
        acc = acc + dI2[k]
    np.testing.assert_allclose(acc, I2, rtol=1e-14)
