"""Pins of the oracle's meshblock sampler and AthenaK ghost fill (CPU only).
Reference: grmhd/athenak.py:105-525 (layout, ghost fill), :639-812 (sampler)."""
import numpy as np
import pytest

from mahakala_amd.grmhd import synthetic as syn
from oracle import mahakala_np as onp
from tests import helpers

RNG = np.random.Generator(np.random.PCG64(4))


def _datasets_from_field(geo, field_fn):
    X = geo["x1v"].astype(np.float64)[:, None, None, :]
    Y = geo["x2v"].astype(np.float64)[:, None, :, None]
    Z = geo["x3v"].astype(np.float64)[:, :, None, None]
    X, Y, Z = np.broadcast_arrays(X, Y, Z)
    fields = np.stack([field_fn(c, X, Y, Z) for c in range(8)]).astype(np.float32)
    ds = {k: geo[k] for k in ("x1v", "x2v", "x3v", "x1f", "x2f", "x3f", "Levels", "LogicalLocations")}
    ds.update(uov=fields[:5], B=fields[5:], VariableNames=list(syn.VARIABLE_NAMES))
    return ds


def _linear(c, X, Y, Z):
    # dyadic coefficients: every value is exact in FP32 and in the averages
    return 1.0 + 0.25 * (c + 1) * X - 0.125 * c * Y + 0.0625 * (c % 3) * Z


def _two_level(n=4):
    keys = [(0, i, j, k) for k in range(2) for j in range(2) for i in range(2) if (i, j, k) != (1, 1, 1)]
    keys += [(1, i, j, k) for k in range(2, 4) for j in range(2, 4) for i in range(2, 4)]
    return syn._geometry(keys, 8.0, 2, n)


def test_same_level_halo_is_neighbour_interior():
    geo = syn.uniform_layout(extent=8.0, nblocks=2, n=4)
    ds = _datasets_from_field(geo, lambda c, X, Y, Z: np.sin(X + c) * np.cos(0.3 * Y) + 0.1 * Z)
    blk = helpers.oracle_blocks(ds)
    idx = {tuple([0] + list(l)): b for b, l in enumerate(geo["LogicalLocations"])}
    b0, b1 = idx[(0, 0, 0, 0)], idx[(0, 1, 0, 0)]
    np.testing.assert_array_equal(blk[b0, :, 1:-1, 1:-1, -1], blk[b1, :, 1:-1, 1:-1, 1])
    np.testing.assert_array_equal(blk[b1, :, 1:-1, 1:-1, 0], blk[b0, :, 1:-1, 1:-1, -2])
    # outer domain boundary stays zero
    assert np.all(blk[b0, :, 1:-1, 1:-1, 0] == 0)


def test_trilinear_is_exact_on_linear_fields():
    geo = syn.uniform_layout(extent=8.0, nblocks=2, n=8)
    ds = _datasets_from_field(geo, _linear)
    snap = helpers.np_snapshot(ds, 0.5)
    S = np.zeros((500, 8))
    S[:, 1:4] = RNG.uniform(-7.0, 7.0, (500, 3))   # away from the domain edge (dx = 1)
    prims, mb = onp.sample_prims(snap, S)
    assert np.all(mb >= 0)
    for c in range(8):
        np.testing.assert_allclose(prims[:, c], _linear(c, S[:, 1], S[:, 2], S[:, 3]), rtol=1e-12, atol=1e-12)


def test_restriction_is_exact_on_linear_fields_faces_and_corners():
    geo = _two_level(n=4)
    ds = _datasets_from_field(geo, _linear)
    blk = helpers.oracle_blocks(ds)
    lv, lg = geo["Levels"], geo["LogicalLocations"]
    b = [m for m in range(len(lv)) if lv[m] == 0 and tuple(lg[m]) == (0, 1, 1)][0]  # x-neighbour of the fine region
    x1v, x2v, x3v = geo["x1v"][b], geo["x2v"][b], geo["x3v"][b]
    dx = float(x1v[1] - x1v[0])
    xh = x1v[-1] + dx  # halo cell centre beyond the +x face
    for kk in range(1, 5):
        for jj in range(1, 5):
            expect = _linear(0, xh, x2v[jj - 1], x3v[kk - 1])
            assert blk[b, 0, kk, jj, -1] == pytest.approx(expect, rel=1e-6)
    # corner halo next to the finer block's corner
    b = [m for m in range(len(lv)) if lv[m] == 0 and tuple(lg[m]) == (0, 0, 0)][0]
    xh, yh, zh = (geo[k][b][-1] + dx for k in ("x1v", "x2v", "x3v"))
    assert blk[b, 0, -1, -1, -1] == pytest.approx(_linear(0, xh, yh, zh), rel=1e-6)


def test_fine_to_coarse_edge_halo_closed_form():
    """Along a fine-to-coarse edge the reference's loop reads fine cell 0 on
    its first iteration and the previous iteration's parity after it, so for
    a field g of the along-edge index only: (g(0) + 4 g(2m) + 3 g(2m+1)) / 8."""
    n = 4
    geo = _two_level(n=n)
    ds = _datasets_from_field(geo, lambda c, X, Y, Z: 1.0 + 0.5 * X)  # varies along x only
    blk = helpers.oracle_blocks(ds)
    lv, lg = geo["Levels"], geo["LogicalLocations"]
    # coarse block (1,0,0): its edge (dj=+1, dk=+1, along x) faces the fine blocks
    b = [m for m in range(len(lv)) if lv[m] == 0 and tuple(lg[m]) == (1, 0, 0)][0]
    fine_dx = float(geo["x1v"][lv == 1][0][1] - geo["x1v"][lv == 1][0][0])
    x0 = float(geo["x1f"][b][0])
    for e in range(2):                       # the two fine children along the edge
        cx0 = x0 + e * n * fine_dx           # first fine cell of child e
        g = lambda q: 1.0 + 0.5 * (cx0 + (q + 0.5) * fine_dx)  # noqa: E731
        for m in range(n // 2):
            expect = (g(0) + 4 * g(2 * m) + 3 * g(2 * m + 1)) / 8
            assert blk[b, 0, -1, -1, 1 + e * n // 2 + m] == pytest.approx(expect, rel=1e-6)


def test_coarse_to_fine_injection():
    geo = _two_level(n=4)
    ds = _datasets_from_field(geo, lambda c, X, Y, Z: np.floor(X) + 10 * np.floor(Y) + 100 * np.floor(Z) + c)
    blk = helpers.oracle_blocks(ds)
    lv, lg = geo["Levels"], geo["LogicalLocations"]
    coarse = [m for m in range(len(lv)) if lv[m] == 0]
    b = [m for m in range(len(lv)) if lv[m] == 1 and tuple(lg[m]) == (2, 2, 2)][0]
    dx = float(geo["x1v"][b][1] - geo["x1v"][b][0])
    # halo cells on the -x face: value of the coarse cell containing the halo centre
    xh = geo["x1v"][b][0] - dx
    for kk in range(1, 5):
        for jj in range(1, 5):
            y, z = geo["x2v"][b][jj - 1], geo["x3v"][b][kk - 1]
            for m in coarse:
                if (geo["x1f"][m][0] < xh <= geo["x1f"][m][-1] and geo["x2f"][m][0] < y <= geo["x2f"][m][-1]
                        and geo["x3f"][m][0] < z <= geo["x3f"][m][-1]):
                    cdx = geo["x1v"][m][1] - geo["x1v"][m][0]
                    i = int((xh - geo["x1f"][m][0]) // cdx)
                    j = int((y - geo["x2f"][m][0]) // cdx)
                    k = int((z - geo["x3f"][m][0]) // cdx)
                    assert blk[b, 3, kk, jj, 0] == ds["uov"][3, m, k, j, i]


def test_block_lookup_brute_force_vs_grid():
    ds = helpers.small_two_level(n=4, seed=1)
    cs = helpers.c_snapshot(ds)
    snap = helpers.np_snapshot(ds, 0.5)
    S = np.zeros((4000, 8))
    S[:, 1:4] = RNG.uniform(-9, 9, (4000, 3))
    faces = np.unique(np.concatenate([ds["x1f"].ravel(), [-8.0, 8.0]]))
    S[:1000, 1] = RNG.choice(faces, 1000)        # points exactly on faces
    S[1000:1500, 2] = RNG.choice(faces, 500)
    _, mb_c = cs.sample_scalars(S, 0.5, with_blocks=True)
    np.testing.assert_array_equal(mb_c, onp.block_index(snap, S))


def test_four_velocity_and_field_invariants():
    ds = helpers.small_two_level(n=4, seed=2)
    snap = helpers.np_snapshot(ds, 0.5)
    S = np.zeros((300, 8))
    S[:, 1:4] = RNG.uniform(-7.5, 7.5, (300, 3))
    S[:, 4] = 1.0
    S[:, 5:] = RNG.normal(size=(300, 3))
    ucon, ucov, bcon, bcov, g = onp.four_vectors(snap, S)
    ok = onp.radius_cal(S, 0.5) > 2.2
    np.testing.assert_allclose(np.einsum("ni,ni->n", ucon, ucov)[ok], -1.0, rtol=0, atol=1e-10)
    bb = np.einsum("ni,ni->n", bcon, bcov)
    np.testing.assert_allclose(np.einsum("ni,ni->n", ucon, bcov)[ok], 0.0, atol=1e-9 * np.sqrt(bb[ok]).max())
    assert np.all(bb[ok] >= -1e-12)


def test_out_of_domain_sample():
    ds = helpers.small_two_level(n=4, seed=2)
    snap = helpers.np_snapshot(ds, 0.5)
    S = np.array([[0.0, 30.0, 1.0, -2.0, 1.0, 0.3, -0.2, 0.1]])
    sc = onp.get_fluid_scalars_from_geodesics(snap, S[None])
    g = onp.metric(S[:, :4], 0.5)
    alpha = float(1 / np.sqrt(-np.linalg.inv(g)[0, 0, 0]))
    assert sc["dens"][0, 0] == 0 and sc["u"][0, 0] == 0 and sc["b"][0, 0] == 0
    assert sc["pitch_angle"][0, 0] == pytest.approx(np.pi / 3)
    # normal observer: u_mu = (-alpha, 0, 0, 0)
    assert sc["kdotu"][0, 0] == pytest.approx(-alpha * S[0, 4], rel=1e-12)
