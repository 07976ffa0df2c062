"""Pins of the oracle's metric, inverse and metric gradient (CPU only).

Reference: geodesics.py:88-104 (metric), :339-347 (LU inverse), :305
(jacfwd).  Each check is a property the mathematics fixes, chosen so that a
plausible slip (sign of a, a dropped term in f or l, a transposed gradient)
fails it; the last test plants such slips in the oracle and shows the checks
catch them.
"""
import numpy as np
import pytest

from oracle import c_oracle
from oracle import mahakala_np as onp

RNG = np.random.Generator(np.random.PCG64(2))
PTS = np.concatenate([np.zeros((64, 1)), RNG.uniform(-25, 25, (64, 3))], 1)
SPINS = (0.0, 0.3, -0.7, 0.9375, 0.999)


def _eta_l_null(x, a):
    f, l1, l2, l3 = onp._ks_f_l(x[..., 1], x[..., 2], x[..., 3], a)
    return -1.0 + l1 * l1 + l2 * l2 + l3 * l3  # eta^{mu nu} l_mu l_nu with l_0 = 1


@pytest.mark.parametrize("a", SPINS)
def test_metric_invariants(a):
    g = onp.metric(PTS, a)
    np.testing.assert_allclose(np.linalg.det(g), -1.0, rtol=0, atol=1e-10)
    gi = onp.imetric(PTS, a)
    np.testing.assert_allclose(g @ gi, np.broadcast_to(np.eye(4), g.shape), atol=1e-12)
    # l is null under eta, hence g^{mu nu} = eta - f l^mu l^nu in closed form
    np.testing.assert_allclose(_eta_l_null(PTS, a), 0.0, atol=1e-13)
    f, l1, l2, l3 = onp._ks_f_l(PTS[:, 1], PTS[:, 2], PTS[:, 3], a)
    lu = np.stack([-np.ones_like(f), l1, l2, l3], -1)
    closed = np.diag([-1.0, 1, 1, 1]) - f[:, None, None] * lu[:, :, None] * lu[:, None, :]
    np.testing.assert_allclose(gi, closed, atol=1e-12)
    # symmetric, static (no t dependence), and equal to the C restatement
    np.testing.assert_array_equal(g, np.swapaxes(g, -1, -2))
    shifted = PTS.copy()
    shifted[:, 0] += 123.0
    np.testing.assert_array_equal(onp.metric(shifted, a), g)
    for k in range(0, 64, 9):
        np.testing.assert_allclose(c_oracle.metric(PTS[k], a), g[k], rtol=1e-14, atol=1e-15)


def test_schwarzschild_limit():
    g = onp.metric(PTS, 0.0)
    R = np.linalg.norm(PTS[:, 1:], axis=1)
    n = PTS[:, 1:] / R[:, None]
    expect = np.diag([-1.0, 1, 1, 1]) + (2 / R)[:, None, None] * np.einsum(
        "ni,nj->nij", np.concatenate([np.ones((64, 1)), n], 1), np.concatenate([np.ones((64, 1)), n], 1))
    np.testing.assert_allclose(g, expect, rtol=1e-13, atol=1e-15)


@pytest.mark.parametrize("a", SPINS)
def test_jacobian_matches_finite_differences(a):
    jg = onp.jac_metric(PTS, a)
    h = 1e-5
    fd = np.stack([(onp.metric(PTS + h * np.eye(4)[s], a) - onp.metric(PTS - h * np.eye(4)[s], a)) / (2 * h)
                   for s in range(4)], -1)
    np.testing.assert_allclose(jg, fd, rtol=1e-6, atol=1e-9)
    assert np.all(jg[..., 0] == 0)  # d/dt g = 0


def test_rhs_three_formulations_agree():
    """NumPy (dual-number jacfwd + LU) vs C (closed-form gradient, full
    contraction): the geodesic acceleration to 1e-13 relative."""
    s = np.concatenate([PTS, RNG.normal(size=(64, 4))], 1)
    for a in SPINS:
        ref = onp.rhs(s, a)
        got = c_oracle.rhs(s, a)
        np.testing.assert_allclose(got, ref, rtol=0, atol=1e-13 * np.abs(ref).max())


def test_gtt_is_boyer_lindquist():
    """g_tt = -(1 - 2 r / Sigma), Sigma = r^2 + a^2 cos^2 theta, with r from
    radius_cal (an independent formula) and cos theta = z / r."""
    for a in SPINS:
        g = onp.metric(PTS, a)
        r = onp.radius_cal(PTS, a)
        sigma = r * r + a * a * (PTS[:, 3] / r) ** 2
        np.testing.assert_allclose(g[:, 0, 0], -(1 - 2 * r / sigma), rtol=1e-12, atol=1e-14)


def _broken(kind):
    def f_l(x1, x2, x3, a):
        aa = a * a
        zz = x3 * x3
        kk = 0.5 * (x1 * x1 + x2 * x2 + zz - aa)
        rr = onp._sqrt(kk * kk + aa * zz) + kk
        r = onp._sqrt(rr)
        if kind == "f_without_a2z2":
            f = (2.0 * rr * r) / (rr * rr)
        else:
            f = (2.0 * rr * r) / (rr * rr + aa * zz)
        sgn = -1.0 if kind == "sign_of_a" else 1.0
        l1 = (r * x1 + sgn * a * x2) / (rr + aa)
        l2 = (r * x2 - a * x1) / (rr + aa)
        l3 = x3 / r
        return f, l1, l2, l3
    return f_l


@pytest.mark.parametrize("kind", ["sign_of_a", "f_without_a2z2"])
def test_planted_mistakes_are_caught(monkeypatch, kind):
    """A sign slip in l or a dropped term in f must fail the pins above."""
    monkeypatch.setattr(onp, "_ks_f_l", _broken(kind))
    a = 0.7
    g = onp.metric(PTS, a)
    det_bad = np.max(np.abs(np.linalg.det(g) + 1.0)) > 1e-6
    null_bad = np.max(np.abs(_eta_l_null(PTS, a))) > 1e-6
    r = onp.radius_cal(PTS, a)
    sigma = r * r + a * a * (PTS[:, 3] / r) ** 2
    gtt_bad = np.max(np.abs(g[:, 0, 0] + (1 - 2 * r / sigma))) > 1e-6
    assert det_bad or null_bad or gtt_bad
