"""GPU parity of the geodesic kernels against the FP64 oracles.

Reference functions: geodesics.py:29-435.  Tolerances: camera and metric
quantities to ~1e-13 relative (rounding of a different but equivalent
expression), trajectories to 1e-8 relative over ~900 RK4 steps, integer
outcomes (frozen row K, trimmed length T, bisection radii) exact.
"""
import os

import numpy as np
import pytest

import mahakala as ma
from mahakala_amd import geodesics as G
from oracle import c_oracle
from oracle import mahakala_np as onp

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.mark.parametrize("a,inc,n", [(0.9, 60.0, 8), (0.5, 30.0, 7), (-0.7, 85.0, 16), (0.0, 0.0, 5)])
def test_camera_grid(a, inc, n):
    gpu = np.asarray(ma.initialize_geodesics_at_camera(a, inc, 1000, -10, 10, n))
    ref = onp.initialize_geodesics_at_camera(a, inc, 1000, -10, 10, n)
    np.testing.assert_allclose(gpu, ref, rtol=1e-13, atol=1e-12, equal_nan=True)
    if n % 2:  # the centre pixel has k = 0 and is NaN in the reference too
        assert np.isnan(gpu[(n * n) // 2, 5:]).all()


def test_camera_equator_and_pixels():
    gpu = np.asarray(ma.initialize_geodesics_at_camera(0.0, 60, 1000, -15, 15, 60, camera_type="Equator"))
    ref = onp.initialize_geodesics_at_camera(0.0, 60, 1000, -15, 15, 60, camera_type="equator")
    np.testing.assert_allclose(gpu, ref, rtol=1e-13, atol=1e-13)
    rad = np.linspace(0.5, 10, 9)
    ang = np.linspace(0, 6, 9)
    sx, sv = G.get_camera_pixel(40, 1000, rad, ang)
    gpu = np.asarray(G.initial_condition(sx, sv, 0.3))
    rx, rv = onp.get_camera_pixel(40, 1000, rad, ang)
    np.testing.assert_allclose(gpu, onp.initial_condition(rx, rv, 0.3), rtol=1e-13, atol=1e-12)


@pytest.mark.parametrize("a", [0.0, 0.5, 0.94, -0.98])
def test_fp64_rhs_and_rk4_step_near_ulp_level(a):
    """The FP64 right-hand side and RK4 step (geodesics.py:294-336) against the
    C restatement's IEEE evaluation of the same closed form, on 4000 states
    from r = 1.6 M to 1200 M: the device's square roots and reciprocals come
    from hardware estimates refined to within 2 ulp.  Measured on MI355X: each
    state's largest deviation is within 8.6e-15 of its largest component for
    the right-hand side (the C and NumPy restatements differ by up to 4e-15 on
    these states) and within 2.3e-16 for the step."""
    rng = np.random.Generator(np.random.PCG64(7))
    n = 4000
    r = np.exp(rng.uniform(np.log(1.6), np.log(1200), n))
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    x = np.zeros((n, 4))
    x[:, 1:] = u * r[:, None]
    s = np.concatenate([x, rng.normal(size=(n, 4))], 1)
    ref = c_oracle.rhs(s, a)
    err = np.abs(np.asarray(G.rhs(s, a)) - ref).max(1) / np.abs(ref).max(1)
    dt = -0.025 * (r - (1 + np.sqrt(1 - a * a)))
    ref4 = c_oracle.rk4(s, dt, a)
    err4 = np.abs(np.asarray(G.RK4_gen(s, dt, a)) - ref4).max(1) / np.abs(ref4).max(1)
    print(f"[rhs a={a}] per-state max rel: rhs {err.max():.2e} (median {np.median(err):.1e}), "
          f"rk4 {err4.max():.2e} (median {np.median(err4):.1e})", flush=True)
    assert err.max() <= 3e-14
    assert err4.max() <= 2e-15


def test_metric_rhs_rk4():
    rng = np.random.Generator(np.random.PCG64(5))
    x = np.zeros((200, 4))
    x[:, 1:] = rng.uniform(-30, 30, (200, 3))
    for a in (0.0, 0.6, -0.95):
        np.testing.assert_allclose(G.metric(x, a), onp.metric(x, a), rtol=1e-13, atol=1e-14)
        np.testing.assert_allclose(G.imetric(x, a), onp.imetric(x, a), rtol=1e-12, atol=1e-13)
        s = np.concatenate([x, rng.normal(size=(200, 4))], 1)
        ref = onp.rhs(s, a)
        np.testing.assert_allclose(G.rhs(s, a), ref, rtol=1e-11, atol=1e-12 * np.abs(ref).max())
        dt = rng.uniform(-1, 0, 200)
        np.testing.assert_allclose(G.RK4_gen(s, dt, a), onp.RK4_gen(s, dt, a), rtol=1e-11, atol=1e-11)
        np.testing.assert_allclose(G.radius_cal(x, a), onp.radius_cal(x, a), rtol=1e-15)


def test_integrator_history_vs_oracle():
    f = np.load(os.path.join(GOLD, "geodesic_6x6.npz"))
    a = float(f["spin"])
    S, dt = G.geodesic_integrator(10000, f["s0"], 40, 1e-4, a)
    assert S.shape == (int(f["T"]), 36, 8)
    np.testing.assert_allclose(np.asarray(S)[f["rows"]], f["S_rows"], rtol=1e-8, atol=1e-8)
    np.testing.assert_allclose(np.asarray(dt)[f["rows"]], f["dt_rows"], rtol=1e-8, atol=1e-10)
    # frozen row per ray = first zero of dt
    K = np.where((np.asarray(dt) == 0).any(0), np.argmax(np.asarray(dt) == 0, 0), 10000)
    np.testing.assert_array_equal(K, f["K"])


def test_integrate_final_64x64_kerr_camera():
    """64x64, a=0.9, i=60, camera at 1000 M; tol = 0.01 r_EH (the '1.01 r_h' stop)."""
    a = 0.9
    rEH = 1 + np.sqrt(1 - a * a)
    s0 = np.asarray(ma.initialize_geodesics_at_camera(a, 60, 1000, -10, 10, 64))
    import ctypes as C
    from mahakala_amd import _lib
    n = s0.shape[0]
    d0 = _lib.DeviceBuffer.from_array(s0)
    dK, dl, ds = _lib.DeviceBuffer(4 * n), _lib.DeviceBuffer(64 * n), _lib.DeviceBuffer(8 * n)
    _lib.check(_lib.lib().mk_integrate_final(d0.ptr, n, 10000, 40.0, 0.01 * rEH, a, dK.ptr, dl.ptr, ds.ptr, None))
    K = dK.to_array((n,), np.int32)
    last = dl.to_array((n, 8))
    Kc, lastc, selc, mg = c_oracle.integrate_final(s0, 10000, 40.0, 0.01 * rEH, a, margins=True)
    mism = np.flatnonzero(K != Kc)
    r_g = onp.radius_cal(last, a)
    r_c = onp.radius_cal(lastc, a)
    sel = ds.to_array((n,))
    print(f"[parity] config 1 64^2: frozen row differs on {mism.size} of {n} rays; termination radius max rel "
          f"{np.max(np.abs(r_g - r_c) / r_c):.2e}; select value max rel {np.max(np.abs(sel - selc) / np.abs(selc)):.2e}")
    detail = [(int(q), int(K[q]), int(Kc[q]), float(mg[q, 0] if K[q] > Kc[q] else mg[q, 1])) for q in mism[:10]]
    assert mism.size == 0, f"frozen rows differ: (ray, K gpu, K c, threshold margin) {detail}"
    # termination radii (north_star: 1e-6 M) and the photon-selection value
    assert np.max(np.abs(r_g - r_c)) < 1e-6
    np.testing.assert_allclose(r_g, r_c, rtol=1e-9)
    np.testing.assert_allclose(sel, selc, rtol=1e-9)
    T = C.c_long()
    _lib.check(_lib.lib().mk_trim_rows(dK.ptr, n, 10000, C.byref(T), None))
    assert T.value == c_oracle.trim_rows(Kc, 10000)


def test_shadow_reference_cases():
    """tests/test_shadows.py of the reference, on the closed-form fixture, plus
    identity with the C restatement (the result is dyadic)."""
    f = np.load(os.path.join(GOLD, "shadow_bardeen.npz"))
    for c in range(4):
        a, inc = float(f[f"case{c}_spin"]), float(f[f"case{c}_inc"])
        ang, target = f[f"case{c}_angles"], f[f"case{c}_radii"]
        radii = ma.find_shadow_bisection_angles(a, inc, ang)
        assert np.allclose(radii, target, rtol=1e-2)
        np.testing.assert_array_equal(radii, c_oracle.find_shadow_bisection_angles(a, inc, ang))


def test_shadow_64_angles_and_select():
    ang, radii = ma.find_shadow_bisection(0.9, 60, 64)
    assert ang.shape == (65,) and radii[0] == radii[-1]
    ref = c_oracle.find_shadow_bisection_angles(0.9, 60, ang[:-1])
    assert np.max(np.abs(radii[:-1] - ref)) <= 1e-6
    f = np.load(os.path.join(GOLD, "shadow_oracle.npz"))
    np.testing.assert_array_equal(ma.find_shadow_bisection_angles(float(f["spin"]), float(f["inc"]), f["angles"]),
                                  f["radii"])
    sel = G.select_photons_integrator(float(f["inc"]), f["angles"], np.full(6, 5.0), float(f["spin"]))
    np.testing.assert_allclose(sel, f["select_r"], rtol=1e-8)


def test_shadow_rounds_of_several_halvings_equal_the_sequential_bisection(monkeypatch):
    """Up to 4096 angles, mk_shadow_bisect first integrates, in one round, every
    midpoint of the first halvings whose bracket meets a band about the shadow
    radius predicted from the critical curve (64 lanes per angle up to 1024
    angles at |a| <= 0.9, 128 above; 32 from 1025 to 4096 angles), replays
    the halvings while the path stays in the band, and runs the rest in tree
    rounds (2^K - 1 midpoints of the next K halvings at once); above 4096
    angles, tree rounds only (K from the angle count), and one lane per angle
    from 32768 angles on.  The radii are the same bits either way: with the
    default band, with a band of 1e-4 M (the path leaves it at most angles, so
    the tree rounds finish them), and with the band off (MK_BISECT_BAND=0,
    tree rounds from the first halving)."""
    ang = np.arange(4096) / 4096 * 2 * np.pi
    seq = ma.find_shadow_bisection_angles(0.9, 60, np.tile(ang, 16))  # 65536 angles: one lane each
    np.testing.assert_array_equal(seq[:4096], seq[4096:8192])
    for band in (None, "1e-4", "0"):
        if band is None:
            monkeypatch.delenv("MK_BISECT_BAND", raising=False)
        else:
            monkeypatch.setenv("MK_BISECT_BAND", band)
        for sub in (ang[::64], ang[::16], ang[::4], ang[::2], ang):  # 64, 256, 1024, 2048, 4096 angles
            got = ma.find_shadow_bisection_angles(0.9, 60, sub)
            np.testing.assert_array_equal(got, seq[:4096][np.isin(ang, sub)],
                                          err_msg=f"band {band}, {sub.size} angles")
    monkeypatch.delenv("MK_BISECT_BAND", raising=False)
    # near-extremal, nearly edge-on (the prediction is 0.03 M off on the flat
    # side) and nearly face-on, against the C restatement
    for a, inc in ((0.99, 85.0), (0.998, 2.0), (-0.95, 70.0)):
        sub = ang[::128]
        np.testing.assert_array_equal(ma.find_shadow_bisection_angles(a, inc, sub),
                                      c_oracle.find_shadow_bisection_angles(a, inc, sub), err_msg=f"a={a} i={inc}")


@pytest.mark.parametrize("dist,err,max_it", [(200.0, 1e-3, 40), (1000.0, 1e-3, 5), (1000.0, 0.1, 40),
                                              (1000.0, 1e-6, 40)])
def test_shadow_bisection_other_distances_and_depths(dist, err, max_it):
    """mk_shadow_bisect at a closer camera (the distant-camera prediction is
    further off) and at 5, 7 and 24 halvings (the band over every level, and a
    band that fits only a narrow path before the tree rounds): the C
    restatement's radii, bit for bit, for 64 angles (one and two waves per
    angle) and 2048 angles (32 lanes per angle)."""
    from mahakala_amd import _lib
    lib = _lib.lib()
    for n, a, inc in ((64, 0.9, 60.0), (64, 0.95, 80.0), (2048, 0.5, 30.0)):
        ang = np.arange(n) / n * 2 * np.pi
        out = np.empty(n)
        _lib.check(lib.mk_shadow_bisect(a, inc, ang.ctypes.data, n, 2000, err, max_it, dist, out.ctypes.data))
        ref = c_oracle.find_shadow_bisection_angles(a, inc, ang, 2000, err, max_it, distance=dist)
        print(f"[shadow] a={a} i={inc} {n} angles, distance {dist}, {c_oracle.bisect_iterations(err, max_it)} "
              f"halvings: identical {np.array_equal(out, ref)}")
        np.testing.assert_array_equal(out, ref, err_msg=f"a={a} i={inc} n={n}")


def test_schwarzschild_shadow():
    r = ma.find_shadow_bisection_angles(0.0, 17.0, np.linspace(0, 6, 12))
    assert np.all(np.abs(r - np.sqrt(27)) < 9.5 / 2 ** 14 + 1e-3)


def test_history_chunked_by_device_budget_equals_one_chunk():
    """geodesic_integrator writes a history larger than its device budget in
    chunks of rays instead of failing (the reference chunks make_image by
    max_chunk_bytes, images.py:62-78; geodesics.py:233-281 returns one
    array): the same S and final_dt bits as one chunk, with the global T."""
    a = 0.9
    s0 = np.asarray(ma.initialize_geodesics_at_camera(a, 60, 1000, -10, 10, 24))
    S1, D1 = G.geodesic_integrator(10000, s0, 40, 1e-4, a, max_chunk_bytes=1 << 40)
    T, n = S1.shape[:2]
    budget = T * G.HISTORY_ROW_BYTES * 100   # 100 rays per chunk: 6 chunks, the last partial
    assert G.history_chunk_rays(n, T, budget) == 100
    S2, D2 = G.geodesic_integrator(10000, s0, 40, 1e-4, a, max_chunk_bytes=budget)
    print(f"[planner] T = {T}, {n} rays, chunks of 100 rays: S and final_dt identical "
          f"{np.array_equal(S1, S2) and np.array_equal(D1, D2)}")
    np.testing.assert_array_equal(S2, S1)
    np.testing.assert_array_equal(D2, D1)
    S3, D3 = G.geodesic_integrator(10000, s0, 40, 1e-4, a)   # default: half of the free device memory
    np.testing.assert_array_equal(S3, S1)
    np.testing.assert_array_equal(D3, D1)
    # one ray per chunk, and the trim rule still global
    S4, D4 = G.geodesic_integrator(10000, s0[:7], 40, 1e-4, a, max_chunk_bytes=1)
    S5, D5 = G.geodesic_integrator(10000, s0[:7], 40, 1e-4, a)
    np.testing.assert_array_equal(S4, S5)
    np.testing.assert_array_equal(D4, D5)
