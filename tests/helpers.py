"""Shared test helpers: build oracle-side views of the synthetic snapshots."""
from __future__ import annotations

import numpy as np

from oracle import c_oracle, mahakala_np as onp

CANON = ("dens", "eint", "velx", "vely", "velz", "bcc1", "bcc2", "bcc3")


def lookup_grid(ds):
    """Leaf-id grid at the finest block size (what the C oracle reads)."""
    lv = np.asarray(ds["Levels"])
    f = [np.asarray(ds[k], dtype=np.float64) for k in ("x1f", "x2f", "x3f")]
    lo = np.stack([a[:, 0] for a in f], 1)
    hi = np.stack([a[:, -1] for a in f], 1)
    b = int(np.argmax(lv))
    w = hi[b] - lo[b]
    g0 = lo.min(0)
    G = np.rint((hi.max(0) - g0) / w).astype(int)
    assert G[0] == G[1] == G[2], "C oracle grid is cubic"
    grid = np.full((G[2], G[1], G[0]), -1, dtype=np.int32)
    for m in range(lo.shape[0]):
        c0 = np.rint((lo[m] - g0) / w).astype(int)
        c1 = np.rint((hi[m] - g0) / w).astype(int)
        grid[c0[2]:c1[2], c0[1]:c1[1], c0[0]:c1[0]] = m
    return grid, g0, float(w[0]), lo, hi


def oracle_blocks(ds):
    """Ghost-padded blocks from the NumPy restatement of the AthenaK loader."""
    return onp.build_meshblocks(ds["uov"], ds["B"], ds["Levels"], ds["LogicalLocations"])


def c_snapshot(ds, blocks_file_order=None):
    """C-oracle snapshot; blocks default to the NumPy oracle's ghost fill."""
    if blocks_file_order is None:
        blocks_file_order = oracle_blocks(ds)
    names = [v.lower() for v in ds["VariableNames"]]
    order = [names.index(c) for c in CANON]
    canon = np.ascontiguousarray(blocks_file_order[:, order].astype(np.float32))
    grid, g0, gh, lo, hi = lookup_grid(ds)
    v = [np.asarray(ds[k], dtype=np.float64) for k in ("x1v", "x2v", "x3v")]
    v0 = np.stack([a[:, 0] for a in v], 1)
    dx = np.stack([a[:, 1] - a[:, 0] for a in v], 1)
    return c_oracle.CSnapshot(canon, lo, hi, v0, dx, grid, g0, gh)


def c_snapshot_from_model(model):
    """C-oracle snapshot holding exactly the GPU's ghost-filled cells."""
    cells = model.device_snapshot.download()          # (nmb, k, j, i, canon)
    canon = np.ascontiguousarray(np.moveaxis(cells, -1, 1))
    ds = dict(Levels=model.Levels, x1f=model.x1f, x2f=model.x2f, x3f=model.x3f)
    grid, g0, gh, lo, hi = lookup_grid(ds)
    v = [np.asarray(a, dtype=np.float64) for a in (model.x1v, model.x2v, model.x3v)]
    v0 = np.stack([a[:, 0] for a in v], 1)
    dx = np.stack([a[:, 1] - a[:, 0] for a in v], 1)
    return c_oracle.CSnapshot(canon, lo, hi, v0, dx, grid, g0, gh)


def np_snapshot(ds, bhspin, fluid_gamma=13. / 9, blocks=None):
    if blocks is None:
        blocks = oracle_blocks(ds)
    return onp.Snapshot(blocks, ds["x1v"], ds["x2v"], ds["x3v"], ds["x1f"], ds["x2f"], ds["x3f"],
                        ds["VariableNames"], bhspin, fluid_gamma)


def small_two_level(n=4, seed=0, bhspin=0.5):
    """Tiny 2-level, 2:1-balanced mesh with smooth random fields (tests)."""
    from mahakala_amd.grmhd import synthetic as syn
    keys = []
    for k in range(2):
        for j in range(2):
            for i in range(2):
                if (i, j, k) == (1, 1, 1):
                    continue
                keys.append((0, i, j, k))
    for k in range(2, 4):
        for j in range(2, 4):
            for i in range(2, 4):
                keys.append((1, i, j, k))
    geo = syn._geometry(keys, 8.0, 2, n)
    rng = np.random.Generator(np.random.PCG64(seed))
    nmb = geo["nmb"]
    coef = rng.normal(size=(8, 4))
    X = geo["x1v"].astype(np.float64)[:, None, None, :]
    Y = geo["x2v"].astype(np.float64)[:, None, :, None]
    Z = geo["x3v"].astype(np.float64)[:, :, None, None]
    X, Y, Z = np.broadcast_arrays(X, Y, Z)
    fields = np.stack([c[0] + c[1] * np.sin(0.3 * X) + c[2] * np.cos(0.2 * Y) + c[3] * 0.05 * Z * X
                       for c in coef]).astype(np.float32)
    fields[0] = np.abs(fields[0]) + 0.5
    fields[4] = np.abs(fields[4]) + 0.1
    ds = {k: geo[k] for k in ("x1v", "x2v", "x3v", "x1f", "x2f", "x3f", "Levels", "LogicalLocations")}
    ds.update(uov=fields[:5], B=fields[5:], VariableNames=list(syn.VARIABLE_NAMES))
    return ds


MSUN = 1.989e33


def image_parity_vs_c(img, steps, cs, a, inc, res, nsteps=10000, nu=230e9, fov=20.0, rtol=1e-4, flux_rtol=1e-5,
                      label="", dist=1000.0):
    """Plain FP64 GPU image and step counts against the C restatement of
    make_image (images.py:30-144): the same accepted-step count on every ray,
    the same finite/NaN pattern, every pixel with |I| above 1e-6 of the peak
    within rtol, and the total flux within flux_rtol (of the summed |I|, so
    that images with negative pixels -- the explicit-Euler transfer of
    transfer.py:106-119 is unstable where alpha |dt| L > 2 -- are held to the
    same bar).  Prints the counts.  A step-count mismatch is reported with the
    relative distance of the deciding candidate from the freeze threshold
    (geodesics.py:249-252), which is rounding-small when that is its cause."""
    ref, rst, mg = cs.render(a, inc, dist, fov, res, nsteps, 13. / 9, 40.0, 1e26, 6.2e9 * MSUN, nu, margins=True)
    img = np.asarray(img).ravel()
    steps = np.asarray(steps).ravel()
    fin = np.isfinite(ref)
    peak = float(np.max(np.abs(ref[fin]))) if fin.any() else 0.0
    mask = fin & (np.abs(ref) > 1e-6 * peak)
    with np.errstate(all="ignore"):
        rel = np.abs(img[mask] - ref[mask]) / np.abs(ref[mask])
    nbad = int(np.count_nonzero(~(rel <= rtol)))
    dK = steps - rst
    mism = np.flatnonzero(dK)
    total = float(np.sum(np.abs(ref[fin])))
    flux = abs(float(np.sum(img[fin])) - float(np.sum(ref[fin]))) / total if total > 0 else 0.0
    print(f"[parity] {label} {res}^2: {int(mask.sum())} pixels checked, {nbad} beyond {rtol:g} "
          f"(max rel {float(rel.max()) if rel.size else 0.0:.2e}), step counts differ on {mism.size} of {steps.size} "
          f"rays, flux rel err {flux:.2e}, negative pixels {int(np.count_nonzero(ref[fin] < 0))}")
    detail = [(int(q), int(steps[q]), int(rst[q]), float(mg[q, 0] if dK[q] > 0 else mg[q, 1])) for q in mism[:10]]
    assert mism.size == 0, f"step counts differ: (ray, K gpu, K c, threshold margin) {detail}"
    np.testing.assert_array_equal(np.isfinite(img), fin)
    assert nbad == 0, f"{nbad} pixels beyond {rtol:g}"
    assert flux <= flux_rtol, f"flux rel err {flux:.3e}"
    if not fin.any() or peak == 0.0:
        assert np.all(img[fin] == 0.0)
    return ref.reshape(res, res)
