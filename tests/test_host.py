"""Host-side logic that needs no GPU: synthetic snapshot layouts, parameter
validation, the bench's one-GPU contract."""
import os
import subprocess
import sys

import numpy as np
import pytest

from mahakala_amd import images
from mahakala_amd.grmhd import synthetic as syn

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _neighbour_levels_ok(geo):
    """Every pair of leaves whose closed boxes touch (face, edge or corner)
    differs by at most one level (the 2:1 balance athenak.py:162 needs)."""
    lo = np.stack([geo[k][:, 0] for k in ("x1f", "x2f", "x3f")], 1).astype(np.float64)
    hi = np.stack([geo[k][:, -1] for k in ("x1f", "x2f", "x3f")], 1).astype(np.float64)
    touch = np.all((lo[:, None, :] <= hi[None, :, :]) & (lo[None, :, :] <= hi[:, None, :]), axis=2)
    dl = np.abs(geo["Levels"][:, None].astype(int) - geo["Levels"][None, :].astype(int))
    return not np.any(touch & (dl > 1))


def test_octree_layout_is_512_leaves_and_2to1_balanced():
    geo = syn.octree_layout()
    assert geo["nmb"] == 512
    assert np.bincount(geo["Levels"]).tolist() == [32, 224, 256]
    assert _neighbour_levels_ok(geo)
    # leaves tile [-50, 50]^3 exactly
    w = geo["x1f"][:, -1].astype(np.float64) - geo["x1f"][:, 0]
    vol = np.sum(w ** 3)
    assert vol == pytest.approx(100.0 ** 3)


def test_uniform_layout_is_256_cubed():
    geo = syn.uniform_layout()
    assert geo["nmb"] == 512 and geo["n"] == 32
    dx = geo["x1v"][:, 1] - geo["x1v"][:, 0]
    assert np.all(dx == np.float32(0.3125))
    assert geo["x1f"].min() == -40 and geo["x1f"].max() == 40


def test_torus_fields_are_physical():
    x, y, z = np.meshgrid(np.linspace(-30, 30, 41), np.linspace(-30, 30, 41), np.linspace(-10, 10, 11))
    f = syn.torus_fields(x, y, z, 0.5)
    assert np.all(f["dens"] >= 0) and np.all(f["u"] >= 0)
    assert np.all(np.isfinite(np.stack(list(f.values()))))
    b2 = f["bcc1"] ** 2 + f["bcc2"] ** 2
    assert np.nanmax(b2[f["dens"] > 0] / f["dens"][f["dens"] > 0]) < 10.0


def test_render_params_validation():
    with pytest.raises(ValueError):
        images.render_params(0.5, precision="fp16")
    p = images.render_params(0.5, precision="mixed-zone")
    assert p.precision == 1 and p.fp64_radius == 4.0 and p.ring_radius == 5.0
    p = images.render_params(0.5, precision="mixed")
    assert p.fp64_radius == 0.0 and p.ring_radius == 0.0


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_history_chunk_planner_arithmetic():
    """Rays per chunk of a materialised history: T rows of 72 bytes per ray
    within the budget, at least one ray, at most all of them."""
    from mahakala_amd import geodesics as G
    assert G.HISTORY_ROW_BYTES == 72
    assert G.history_chunk_rays(576, 1000, 72 * 1000 * 100) == 100
    assert G.history_chunk_rays(576, 1000, 72 * 1000 * 100 - 1) == 99
    assert G.history_chunk_rays(576, 1000, 1) == 1
    assert G.history_chunk_rays(576, 1000, 1 << 50) == 576
    assert G.history_chunk_rays(576, 0, 72 * 10) == 10
