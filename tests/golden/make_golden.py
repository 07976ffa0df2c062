"""Writes the golden fixtures in tests/golden/ from the NumPy oracle.

Run from the repository root:  python tests/golden/make_golden.py
Every fixture is data only (inputs and the oracle's outputs).  The shadow
fixture replaces the reference's pickled tests/data/shadow_data.npy, which
numpy.load refuses without allow_pickle: it holds the same four (spin,
inclination, angle-count) cases with radii from the Bardeen closed form,
which is what the reference fixture matches.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import mahakala_np as onp  # noqa: E402
from tests import helpers  # noqa: E402

# (spin, inclination, angles) of the reference test (tests/test_shadows.py:28-45)
SHADOW_CASES = [(0.9, 90.0, 100), (0.2, 90.0, 100), (0.44, 52.0, 80), (0.71, 13.0, 25)]


def shadow_fixture():
    out = {}
    for c, (a, inc, n) in enumerate(SHADOW_CASES):
        rp, rm = onp.photon_orbit_radii(a)
        r = np.linspace(rp, rm, 200001)[1:-1]
        ang, rad = onp.bardeen_curve(a, inc, r)
        # n points evenly spaced along the valid r range, angles from pi to 0
        idx = np.linspace(0, ang.size - 1, n).round().astype(int)
        out[f"case{c}_spin"] = a
        out[f"case{c}_inc"] = inc
        out[f"case{c}_angles"] = ang[idx]
        out[f"case{c}_radii"] = rad[idx]
    np.savez(os.path.join(HERE, "shadow_bardeen.npz"), **out)


def geodesic_fixture():
    a, inc = 0.9, 60.0
    s0 = onp.initialize_geodesics_at_camera(a, inc, 1000, -10, 10, 6)
    S, D, K = onp.geodesic_integrator(10000, s0, 40, 1e-4, a, return_steps=True)
    rows = np.array([0, 1, 2, 50, 200, S.shape[0] - 2, S.shape[0] - 1])
    np.savez(os.path.join(HERE, "geodesic_6x6.npz"), spin=a, inc=inc, s0=s0, K=K, T=S.shape[0], rows=rows,
             S_rows=S[rows], dt_rows=D[rows], S_last=S[-1], dt_sum=D.sum(0))
    ang = np.linspace(0.1, 3.0, 6)
    radii = onp.find_shadow_bisection_angles(a, inc, ang)
    sel = onp.select_photons_integrator(inc, ang, np.full(6, 5.0), a)
    np.savez(os.path.join(HERE, "shadow_oracle.npz"), spin=a, inc=inc, angles=ang, radii=radii, select_r=sel)


def transfer_fixture():
    rng = np.random.Generator(np.random.PCG64(7))
    n = 64
    Ne = 10 ** rng.uniform(3, 7, n)
    Th = 10 ** rng.uniform(-1, 2, n)
    B = 10 ** rng.uniform(-1, 2, n)
    pitch = rng.uniform(0.01, np.pi - 0.01, n)
    nu = 10 ** rng.uniform(10, 13, n)
    Th[:4] = [0.2, 0.31, 100., 1.]            # cut branches
    nu[4] = 1e30                               # X > 1e12
    j, al = onp.synchrotron_coefficients(Ne, Th, B, pitch, nu, invariant=True, rescale_nu=1 / 230e9)
    jr, alr = onp.synchrotron_coefficients(Ne, Th, B, pitch, nu, invariant=False)
    T, npx = 40, 16
    jj = 10 ** rng.uniform(-16, -13, (T, npx))
    aa = 10 ** rng.uniform(-18, -15, (T, npx))
    dt = -10 ** rng.uniform(-3, 0, (T, npx))
    dt[-5:, :4] = 0.0
    L = 9.2e14
    I, dI = onp.solve_specific_intensity(jj, aa, dt, L, dIs=True)
    att = onp.solve_attenuated_emissivity(jj, aa, dt, L)
    dens = 10 ** rng.uniform(-3, 0, n)
    u = 10 ** rng.uniform(-4, -1, n)
    beta = 10 ** rng.uniform(-1, 2, n)
    th = onp.rlow_rhigh_model(dens, u, beta, r_high=40)
    np.savez(os.path.join(HERE, "transfer.npz"), Ne=Ne, Th=Th, B=B, pitch=pitch, nu=nu, j=j, al=al, j_raw=jr,
             al_raw=alr, jj=jj, aa=aa, dt=dt, L=L, I=I, dI=dI, att=att, dens=dens, u=u, beta=beta, theta=th)


def snapshot_fixture():
    ds = helpers.small_two_level(n=4, seed=3)
    blocks = helpers.oracle_blocks(ds)
    # the device holds FP32 cells: sample the FP32-rounded halo averages
    snap = helpers.np_snapshot(ds, 0.5, blocks=blocks.astype(np.float32).astype(np.float64))
    rng = np.random.Generator(np.random.PCG64(11))
    m = 400
    S = np.zeros((m, 8))
    S[:, 1:4] = rng.uniform(-9, 9, (m, 3))
    S[:5, 1:4] = [[0.0, 0.0, 0.0], [8.0, 8.0, 8.0], [-8.0, 1.0, 1.0], [4.0, 4.0, 4.0], [1e-9, 2.0, 3.0]]
    S[:, 4] = 1.0
    S[:, 5:8] = rng.normal(size=(m, 3))
    sc = onp.get_fluid_scalars_from_geodesics(snap, S[None])
    mb = onp.block_index(snap, S)
    prims, _ = onp.sample_prims(snap, S)
    np.savez(os.path.join(HERE, "snapshot_2level.npz"), blocks=blocks.astype(np.float32), S=S, mb=mb,
             prims=prims, **{k: v[0] for k, v in sc.items()})


def image_fixture():
    ds = helpers.small_two_level(n=4, seed=3)
    blocks = helpers.oracle_blocks(ds).astype(np.float32).astype(np.float64)
    snap = helpers.np_snapshot(ds, 0.5, blocks=blocks)
    img = onp.make_image(snap, camera_inclination=50, fov=16, resolution=10, mass_scale=1e24,
                         observing_frequency=230e9)
    np.savez(os.path.join(HERE, "image_2level_10x10.npz"), image=img, inc=50.0, fov=16.0, res=10,
             mass_scale=1e24, nu=230e9, spin=0.5)


if __name__ == "__main__":
    which = sys.argv[1:] or ["shadow", "geodesic", "transfer", "snapshot", "image"]
    for w in which:
        globals()[w + "_fixture"]()
        print("wrote", w)
