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
