"""GPU parity of the fused image kernel (images.py:30-144 of the reference).

Plain FP64 against the FP64 restatements: pixel intensities within 1e-4
relative (pixels above 1e-6 of the peak) and total flux within 1e-5
relative.  Mixed FP32/FP64 against plain FP64: total flux within 1e-3.
"""
import os

import numpy as np
import pytest

from mahakala_amd import images
from mahakala_amd.grmhd import synthetic as syn
from mahakala_amd.grmhd.athenak import AthenakFluidModel
from tests import helpers

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
MSUN = 1.989e33
MK_RENDER_STATIC_TILES = 1024  # include/mahakala_hip.h


def _pixel_close(img, ref, rtol=1e-4):
    mask = ref > 1e-6 * ref.max()
    rel = np.abs(img[mask] - ref[mask]) / ref[mask]
    return rel, mask


@pytest.fixture(scope="module")
def torus():
    ds = syn.uniform_torus_snapshot(0.5, seed=0)
    return AthenakFluidModel.from_datasets(ds, 0.5, fluid_gamma=13. / 9)


def test_image_fixture_from_numpy_oracle():
    """NumPy restatement of make_image (materialised history, backward scan)."""
    f = np.load(os.path.join(GOLD, "image_2level_10x10.npz"))
    ds = helpers.small_two_level(n=4, seed=3)
    m = AthenakFluidModel.from_datasets(ds, float(f["spin"]), fluid_gamma=13. / 9)
    img = images.make_image(m, camera_inclination=float(f["inc"]), fov=float(f["fov"]), resolution=int(f["res"]),
                            mass_scale=float(f["mass_scale"]), observing_frequency=float(f["nu"]))
    ref = f["image"]
    rel, mask = _pixel_close(img, ref)
    assert mask.sum() > 10 and rel.max() < 1e-4
    assert abs(img.sum() - ref.sum()) / ref.sum() < 1e-5
    # the unfused pipeline (GPU integrator + GPU sampler + GPU transfer) agrees too
    gen = images.make_image(_Unfused(m), camera_inclination=float(f["inc"]), fov=float(f["fov"]),
                            resolution=int(f["res"]), mass_scale=float(f["mass_scale"]),
                            observing_frequency=float(f["nu"]))
    rel, _ = _pixel_close(gen, ref)
    assert rel.max() < 1e-4
    # chunked by max_chunk_bytes (the reference's rule, images.py:62-69): 13
    # pixels per chunk here, the same image bit for bit (frozen rows add 0)
    chunked = images.make_image(_Unfused(m), camera_inclination=float(f["inc"]), fov=float(f["fov"]),
                                resolution=int(f["res"]), mass_scale=float(f["mass_scale"]),
                                observing_frequency=float(f["nu"]), max_chunk_bytes=13 * 4 * 20 * 10000)
    np.testing.assert_array_equal(chunked, gen)


class _Unfused:
    """Duck-typed fluid model without a resident snapshot handle."""

    def __init__(self, m):
        self._m = m
        self.bhspin = m.bhspin
        self.fluid_gamma = m.fluid_gamma

    def get_units(self, M, ms):
        return self._m.get_units(M, ms)

    def get_fluid_scalars_from_geodesics(self, S):
        return self._m.get_fluid_scalars_from_geodesics(S)


@pytest.mark.parametrize("res", [48, 256])
def test_uniform_torus_plain_fp64_vs_c_restatement(torus, res):
    """Config 2 (a = 0.5, i = 30 deg, 230 GHz) at 48^2 and at its full 256^2:
    every ray's step count identical and every pixel within 1e-4."""
    p = images.render_params(0.5, camera_inclination=30, fov=20, resolution=res)
    img, steps = images.render(torus.device_snapshot, p, return_steps=True)
    cs = helpers.c_snapshot_from_model(torus)
    ref = helpers.image_parity_vs_c(img, steps, cs, 0.5, 30, res, label="torus a=0.5 i=30")
    assert ref.sum() > 0


@pytest.mark.parametrize("a,inc", [(0.998, 2.0), (-0.7, 88.0), (0.0, 45.0)])
def test_extreme_views_plain_fp64_vs_c_restatement(torus, a, inc):
    """Near face-on at spin 0.998, near edge-on and retrograde, and
    Schwarzschild, 64^2 through the config-2 torus (the request's spin sets
    the geodesics): every step count identical, every pixel within 1e-4.
    The mixed image's flux stays within 1e-3 of plain FP64."""
    res = 64
    p = images.render_params(a, camera_inclination=inc, fov=20, resolution=res)
    img, steps = images.render(torus.device_snapshot, p, return_steps=True)
    cs = helpers.c_snapshot_from_model(torus)
    ref = helpers.image_parity_vs_c(img, steps, cs, a, inc, res, label=f"torus a={a} i={inc}")
    assert ref.sum() > 0
    mixed = images.render(torus.device_snapshot, images.render_params(a, camera_inclination=inc, fov=20,
                                                                      resolution=res, precision="mixed"))
    assert abs(mixed.sum() - img.sum()) <= 1e-3 * img.sum()


def test_random_views_plain_fp64_vs_c_restatement(torus):
    """Six views drawn at random (seed 11): spin in (-0.99, 0.99), inclination
    in (1, 179) deg, fov 8-40 M, camera at 200-1000 M, 86-690 GHz, 40^2
    through the config-2 torus.  Plain FP64 against the C restatement: every
    step count identical, every pixel within 1e-4."""
    rng = np.random.Generator(np.random.PCG64(11))
    cs = helpers.c_snapshot_from_model(torus)
    res = 40
    for k in range(6):
        a, inc = float(rng.uniform(-0.99, 0.99)), float(rng.uniform(1, 179))
        fov, dist = float(rng.uniform(8, 40)), float(rng.choice([200.0, 500.0, 1000.0]))
        nu = float(rng.choice([86e9, 230e9, 345e9, 690e9]))
        p = images.render_params(a, camera_inclination=inc, camera_distance=dist, fov=fov, resolution=res,
                                 observing_frequency=nu)
        img, steps = images.render(torus.device_snapshot, p, return_steps=True)
        helpers.image_parity_vs_c(img, steps, cs, a, inc, res, nu=nu, fov=fov, dist=dist,
                                  label=f"a={a:.3f} i={inc:.1f} fov={fov:.1f} d={dist:g} nu={nu:.3g}")


def test_target_view_plain_fp64_vs_c_restatement():
    """The precision target's view at full size: 256^2, a = 0.9, i = 60 deg,
    the torus generated at a = 0.9 on 512 blocks of 32^3."""
    m9 = AthenakFluidModel.from_datasets(syn.uniform_torus_snapshot(0.9, seed=0), 0.9, fluid_gamma=13. / 9)
    p = images.render_params(0.9, camera_inclination=60, fov=20, resolution=256)
    img, steps = images.render(m9.device_snapshot, p, return_steps=True)
    helpers.image_parity_vs_c(img, steps, helpers.c_snapshot_from_model(m9), 0.9, 60, 256, label="torus a=0.9 i=60")


@pytest.mark.parametrize("precision", ["mixed", "mixed-zone"])
def test_uniform_torus_mixed_vs_plain(torus, precision):
    """Config 2 at its full 256^2: total flux of each mixed preset within 1e-3
    of plain FP64 (the north star's bar; measured ~1.3e-7)."""
    res = 256
    plain = images.render(torus.device_snapshot, images.render_params(0.5, camera_inclination=30, resolution=res))
    mixed = images.render(torus.device_snapshot,
                          images.render_params(0.5, camera_inclination=30, resolution=res, precision=precision))
    err = abs(mixed.sum() - plain.sum()) / plain.sum()
    print(f"[render] 256^2 torus {precision}: flux rel err {err:.3e}")
    assert err < 1e-3


def test_make_image_fused_equals_render(torus):
    img = images.make_image(torus, camera_inclination=30, resolution=24)
    p = images.render_params(0.5, camera_inclination=30, resolution=24)
    np.testing.assert_array_equal(img, images.render(torus.device_snapshot, p))


@pytest.fixture(scope="module")
def octree():
    """3-level 512-leaf tree over +-50 M with log-normal turbulence (a = 0.94)."""
    ds = syn.octree_turbulent_snapshot(0.94, seed=0)
    return AthenakFluidModel.from_datasets(ds, 0.94, fluid_gamma=13. / 9)


def test_octree_turbulence_plain_fp64_vs_c_restatement(octree):
    """a = 0.94, i = 17 deg on the refined tree: block lookup across three
    levels, halos from coarser and finer neighbours, and the turbulent field."""
    res = 40
    p = images.render_params(0.94, camera_inclination=17, fov=20, resolution=res)
    img, steps = images.render(octree.device_snapshot, p, return_steps=True)
    cs = helpers.c_snapshot_from_model(octree)
    ref = helpers.image_parity_vs_c(img, steps, cs, 0.94, 17, res, label="octree a=0.94 i=17")
    assert ref.sum() > 0 and np.all(ref >= 0)


@pytest.mark.parametrize("precision", ["mixed-zone", "mixed"])
def test_octree_1024_mixed_vs_plain_flux(octree, precision):
    """1024^2 rays, a = 0.94, i = 17 deg, 230 GHz: total flux of each mixed
    preset within 1e-3 of plain FP64."""
    res = 1024
    plain = images.render(octree.device_snapshot, images.render_params(0.94, camera_inclination=17, resolution=res))
    mixed = images.render(octree.device_snapshot, images.render_params(0.94, camera_inclination=17, resolution=res,
                                                                        precision=precision))
    assert plain.sum() > 0 and np.all(plain >= 0)
    err = abs(mixed.sum() - plain.sum()) / plain.sum()
    print(f"[render] 1024^2 octree {precision}: flux rel err {err:.3e}")
    assert err < 1e-3
    # the classic form's tiles from a counter against one wave per tile
    p = images.render_params(0.94, camera_inclination=17, resolution=res, precision=precision)
    assert _form_of(octree, p) == "classic"
    p.flags |= MK_RENDER_STATIC_TILES
    np.testing.assert_array_equal(images.render(octree.device_snapshot, p), mixed)


@pytest.mark.parametrize("res,nsteps", [(13, 10000), (37, 10000), (24, 300), (9, 1)])
def test_edge_resolutions_and_step_caps_vs_c_restatement(torus, res, nsteps):
    """Odd resolutions (the centre pixel's ray is NaN, as the reference's),
    resolutions that leave partial 8x8 tiles, and step caps that stop rays
    mid-flight (row N is never emitted) against the C restatement."""
    p = images.render_params(0.5, camera_inclination=30, fov=20, resolution=res, max_nsteps=nsteps)
    img, steps = images.render(torus.device_snapshot, p, return_steps=True)
    cs = helpers.c_snapshot_from_model(torus)
    helpers.image_parity_vs_c(img, steps, cs, 0.5, 30, res, nsteps=nsteps, label=f"edge N={nsteps}")


def test_image_outside_the_domain_is_zero(torus):
    """A field of view that never meets the snapshot: every sample is outside
    every block, and the image is exactly zero."""
    p = images.render_params(0.5, camera_inclination=30, fov=2000, resolution=16)
    img = images.render(torus.device_snapshot, p)
    corners = img[[0, 0, -1, -1], [0, -1, 0, -1]]
    assert np.all(corners == 0.0)
    cs = helpers.c_snapshot_from_model(torus)
    ref, _ = cs.render(0.5, 30, 1000.0, 2000.0, 16, 10000, 13. / 9, 40.0, 1e26, 6.2e9 * MSUN, 230e9)
    np.testing.assert_allclose(img.ravel(), ref, rtol=1e-4, atol=0)


def _form_of(model, params):
    import ctypes as C
    from mahakala_amd import _lib
    f = C.c_int32()
    _lib.check(_lib.lib().mk_render_form(model.device_snapshot.handle, params, C.byref(f)))
    return {v: k for k, v in images.FORMS.items() if k}[f.value]


@pytest.mark.parametrize("precision", ["plain64", "mixed", "mixed-zone"])
@pytest.mark.parametrize("res", [24, 36, 256, 320])
def test_kernel_forms_agree_bit_for_bit(torus, precision, res):
    """The classic form (one wave per tile, tiles taken from a counter or one
    launch slot each), the paired form (geodesic and
    transfer waves, samples handed over through LDS) and the triple form
    (geodesic, gather and emission waves, two LDS rings) evaluate the same
    expressions per ray in the same order: same image and same step counts,
    bit for bit, in every precision.  On a 256-CU GPU the triple form runs
    256^2 as one wave of workgroups (all strided) and 320^2 as two (the
    resident ones strided, the rest in order)."""
    out = {}
    for form in ("classic", "paired", "triple"):
        p = images.render_params(0.5, camera_inclination=30, resolution=res, precision=precision, form=form)
        assert _form_of(torus, p) == form
        out[form] = images.render(torus.device_snapshot, p, return_steps=True)
    # the classic form with one wave per tile instead of tiles from a counter
    p = images.render_params(0.5, camera_inclination=30, resolution=res, precision=precision, form="classic")
    p.flags |= MK_RENDER_STATIC_TILES
    out["classic-static"] = images.render(torus.device_snapshot, p, return_steps=True)
    a, sa = out["classic"]
    assert np.all(np.isfinite(a)) and a.max() > 0
    for form in ("paired", "triple", "classic-static"):
        b, sb = out[form]
        d = np.abs(a - b)
        print(f"[render] {precision} {res}^2 {form} vs classic: max abs diff {d.max():.2e}, "
              f"steps differ {int((sa != sb).sum())}")
        np.testing.assert_array_equal(sa, sb)
        np.testing.assert_array_equal(a, b)


def test_kernel_forms_agree_at_random_views(torus):
    """The four launch forms (classic from a counter, classic one slot per
    tile, paired, triple) at four views drawn at random (seed 5: spin, an
    inclination in (1, 179) deg, fov, frequency) and sizes of 40-72^2 (partial
    tiles included), in plain FP64, mixed and mixed-zone, whose FP64 zone
    decisions then fall where the view puts them: same image and step counts,
    bit for bit."""
    rng = np.random.Generator(np.random.PCG64(5))
    for k in range(4):
        a, inc = float(rng.uniform(-0.98, 0.98)), float(rng.uniform(1, 179))
        fov, nu, res = float(rng.uniform(10, 30)), float(rng.choice([86e9, 230e9, 690e9])), int(rng.integers(40, 73))
        for precision in ("plain64", "mixed", "mixed-zone"):
            out = {}
            for form in ("classic", "paired", "triple", "classic-static"):
                p = images.render_params(a, camera_inclination=inc, fov=fov, resolution=res, observing_frequency=nu,
                                         precision=precision, form=form.split("-")[0])
                if form == "classic-static":
                    p.flags |= MK_RENDER_STATIC_TILES
                out[form] = images.render(torus.device_snapshot, p, return_steps=True)
            base, sb = out["classic"]
            for form in ("paired", "triple", "classic-static"):
                img, st = out[form]
                np.testing.assert_array_equal(st, sb, err_msg=f"{form} {precision} view {k}")
                np.testing.assert_array_equal(img, base, err_msg=f"{form} {precision} view {k}")
        print(f"[render] random view {k} (a={a:.3f} i={inc:.1f} fov={fov:.1f} nu={nu:.3g} {res}^2): "
              f"four forms identical in three precisions")


def test_classic_tile_counter_starts_from_zero_every_launch(torus):
    """The classic form's waves take tiles from a per-stream counter that the
    last wave of each launch resets: back-to-back launches on the null stream
    and on a created stream (a counter of its own) all give the full image."""
    import ctypes as C
    from mahakala_amd import _lib
    res = 512
    p = images.render_params(0.5, camera_inclination=30, resolution=res, precision="mixed")
    assert _form_of(torus, p) == "classic"
    lib = _lib.lib()
    st = C.c_void_p()
    _lib.check(lib.mk_stream_create(C.byref(st)))
    try:
        ref = None
        for k, stream in enumerate([None, None, st.value, st.value, None]):
            buf = _lib.DeviceBuffer(8 * res * res)
            _lib.check(lib.mk_memset(buf.ptr, 0, 8 * res * res, stream))
            _lib.check(lib.mk_render(torus.device_snapshot.handle, p, buf.ptr, None, stream))
            _lib.check(lib.mk_stream_sync(stream))
            img = buf.to_array((res, res))
            print(f"[render] classic launch {k} on {'null' if stream is None else 'created'} stream: "
                  f"{np.count_nonzero(img)} nonzero pixels")
            if ref is None:
                ref = img
                assert np.count_nonzero(img) > res * res // 4
            np.testing.assert_array_equal(img, ref)
    finally:
        _lib.check(lib.mk_stream_destroy(st.value))


@pytest.fixture(scope="module")
def torus_many_blocks():
    """The torus as 4096 blocks of 8^3 cells: block tables of 393 KB, more than
    the LDS budget, so the kernels read them from global memory."""
    return AthenakFluidModel.from_datasets(syn.uniform_torus_snapshot(0.5, seed=0, nblocks=16, n=8), 0.5,
                                           fluid_gamma=13. / 9)


@pytest.mark.parametrize("precision", ["plain64", "mixed"])
def test_triple_form_with_tables_in_global_memory(torus_many_blocks, precision):
    # 36^2: edge tiles with idle lanes; 128^2: full tiles
    """Tables too large for LDS: the triple form reads them from global memory
    and its gather wave keeps each ray's last block record; image and step
    counts equal the classic form's bit for bit, and the form is the default
    for a 128^2 image."""
    for res in (36, 128):
        out = {}
        for form in ("classic", "triple"):
            p = images.render_params(0.5, camera_inclination=30, resolution=res, precision=precision, form=form)
            assert _form_of(torus_many_blocks, p) == form
            out[form] = images.render(torus_many_blocks.device_snapshot, p, return_steps=True)
        (a, sa), (b, sb) = out["classic"], out["triple"]
        assert a.max() > 0 and np.all(np.isfinite(b))
        np.testing.assert_array_equal(sa, sb)
        np.testing.assert_array_equal(a, b)
    p = images.render_params(0.5, camera_inclination=30, resolution=128, precision=precision)
    assert _form_of(torus_many_blocks, p) == "triple"


@pytest.mark.parametrize("form", ["paired", "triple"])
def test_bounded_wait_failure_path(torus, form):
    """With the wait-test flag every wait between the waves of a tile gives up
    after one poll: the launch still completes, a tile whose wave gave up has
    NaN pixels (and step counts -1 when its geodesic wave gave up), every other
    pixel equals the normal render, and the next normal render is unaffected."""
    p = images.render_params(0.5, camera_inclination=30, resolution=64, precision="mixed", form=form)
    ref, sref = images.render(torus.device_snapshot, p, return_steps=True)
    p.flags |= 256  # MK_RENDER_WAIT_TEST
    img, st = images.render(torus.device_snapshot, p, return_steps=True)
    bad = np.isnan(img)
    print(f"[render] {form} wait test: {int(bad.sum())} NaN pixels, {int((st == -1).sum())} failed rays")
    assert bad.any()
    assert np.all(bad[st == -1])
    np.testing.assert_array_equal(img[~bad], ref[~bad])
    ok = st != -1
    np.testing.assert_array_equal(st[ok], sref[ok])
    p.flags &= ~256
    again, sagain = images.render(torus.device_snapshot, p, return_steps=True)
    np.testing.assert_array_equal(again, ref)
    np.testing.assert_array_equal(sagain, sref)


@pytest.mark.parametrize("form", ["classic", "triple"])
@pytest.mark.parametrize("spin,inc", [(0.5, 30.0), (0.98, 85.0)])
def test_tile_order_changes_only_the_schedule(torus, form, spin, inc):
    """Tiles handed out photon ring first (the default; the critical curve at
    the image's spin and inclination) or row-major (MK_RENDER_ROW_MAJOR): the
    same image and step counts, bit for bit, on a launch of several rounds
    whose edge tiles are partly empty (1000^2)."""
    out = {}
    for flag in (0, 512):
        p = images.render_params(spin, camera_inclination=inc, resolution=1000, precision="mixed", form=form)
        p.flags |= flag
        out[flag] = images.render(torus.device_snapshot, p, return_steps=True)
    (a, sa), (b, sb) = out[0], out[512]
    assert a.max() > 0 and np.all(np.isfinite(a))
    np.testing.assert_array_equal(sa, sb)
    np.testing.assert_array_equal(a, b)


def test_form_choice(torus):
    """Launches of up to 15.5 tiles per CU in mixed precision (496^2 on 256
    CUs) and 11 in FP64 (424^2) run the triple form, larger ones the classic
    form; a forced form is honoured."""
    for prec in ("plain64", "mixed"):
        assert _form_of(torus, images.render_params(0.5, resolution=256, precision=prec)) == "triple"
        assert _form_of(torus, images.render_params(0.5, resolution=384, precision=prec)) == "triple"
        assert _form_of(torus, images.render_params(0.5, resolution=448, precision=prec)) == (
            "triple" if prec == "mixed" else "classic")
        assert _form_of(torus, images.render_params(0.5, resolution=512, precision=prec)) == "classic"
        assert _form_of(torus, images.render_params(0.5, resolution=1024, precision=prec)) == "classic"
        assert _form_of(torus, images.render_params(0.5, resolution=1024, precision=prec, form="paired")) == "paired"


@pytest.mark.parametrize("precision", ["plain64", "mixed"])
def test_face_on_image_is_invariant_under_quarter_turns(torus, precision):
    """At inclination 0 the camera looks down the spin axis; the torus, its
    flow and field, the Cartesian block grid and the pixel grid are all
    invariant under a quarter turn about that axis, so the full 256^2 image is
    too (up to the rounding of different arithmetic orders)."""
    p = images.render_params(0.5, camera_inclination=0, fov=20, resolution=256, precision=precision)
    img = images.render(torus.device_snapshot, p)
    assert img.max() > 0
    live = img > 1e-6 * img.max()
    for k in (1, 2, 3):
        rot = np.rot90(img, k)
        rel = np.abs(rot - img)[live] / img[live]
        print(f"[render] {precision} face-on quarter turn {k}: max rel {rel.max():.2e}")
        assert np.median(rel) < (1e-10 if precision == "plain64" else 1e-5)
        assert abs(rot.sum() - img.sum()) <= (1e-10 if precision == "plain64" else 1e-6) * img.sum()
