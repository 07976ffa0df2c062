"""GPU tests of the request server (continuous batching of image and shadow
requests; images.py:30-144 and geodesics.py:405-435 of the reference).

* a request's result is bit-identical alone and mixed with other requests
  (different resolution, spin, inclination, frequency, snapshot, kind), with
  slots reused across launches;
* served images equal mk_render's and served bisections equal
  mk_shadow_bisect's (same arithmetic per ray);
* admission respects the HBM budget (back-pressure, never-fitting requests
  rejected) and wait() times out;
* the persistent form gives the same bits, and its launch ends when the
  server is idle and resumes its ticket count on the next request.
"""
import time

import numpy as np
import pytest

from mahakala_amd import _lib, images
from mahakala_amd.geodesics import find_shadow_bisection_angles
from mahakala_amd.grmhd import synthetic as syn
from mahakala_amd.grmhd.athenak import AthenakFluidModel
from mahakala_amd.serving import Server
from oracle import c_oracle
from tests import helpers

pytestmark = pytest.mark.gpu
MSUN = 1.989e33

# (resolution, spin, inclination, frequency, snapshot index)
# 36^2 is not a multiple of 8: its items run in rows of 64 pixels with a partial last tile
MIX = [(48, 0.5, 30.0, 230e9, 0), (32, -0.7, 75.0, 86e9, 1), (64, 0.9, 10.0, 345e9, 0), (36, 0.0, 50.0, 230e9, 1)]


@pytest.fixture(scope="module")
def models():
    torus = AthenakFluidModel.from_datasets(syn.uniform_torus_snapshot(0.5, seed=0), 0.5, fluid_gamma=13. / 9)
    small = AthenakFluidModel.from_datasets(helpers.small_two_level(n=4, seed=3), 0.5, fluid_gamma=13. / 9)
    return torus, small


def _img_kw(res, spin, inc, nu):
    return dict(bhspin=spin, inclination=inc, resolution=res, observing_frequency=nu, fov=20, mass_scale=1e26,
                max_nsteps=4000)


def _render(model, res, spin, inc, nu, precision):
    p = images.render_params(spin, inc, 1000.0, 1e26, 6.2e9 * MSUN, 40, nu, 20, res, 4000, 13. / 9, precision)
    return images.render(model.device_snapshot, p)


@pytest.mark.parametrize("persistent", [False, True])
@pytest.mark.parametrize("precision", ["plain64", "mixed"])
def test_served_images_equal_render_and_do_not_depend_on_the_mix(models, precision, persistent):
    print(f"[serving] precision={precision} persistent={persistent}", flush=True)
    alone = {}
    # few lanes and short launches: rays of many requests share waves and
    # each slot is reused many times (the persistent form: one launch whose
    # waves take tiles of every request as they are published)
    kw = dict(lanes=2048, steps_per_launch=64, persistent=persistent, image_in_launch=True)
    with Server(precision, **kw) as srv:
        sids = [srv.add_snapshot(m) for m in models]
        for k, (res, spin, inc, nu, s) in enumerate(MIX):
            rid = srv.submit_image(sids[s], **_img_kw(res, spin, inc, nu))
            alone[k] = srv.result(rid, timeout=120)
    with Server(precision, **kw) as srv:
        sids = [srv.add_snapshot(m) for m in models]
        ang = np.linspace(0, 2 * np.pi, 24, endpoint=False)
        sh = srv.submit_shadow(0.3, 40.0, ang, max_steps=1500)
        rids = [srv.submit_image(sids[s], **_img_kw(res, spin, inc, nu)) for res, spin, inc, nu, s in MIX[::-1]]
        mixed = {len(MIX) - 1 - i: srv.result(r, timeout=120) for i, r in enumerate(rids)}
        srv.result(sh, timeout=120)
        st = srv.stats()
    assert st["completed"] == len(MIX) + 1 and st["in_flight"] == 0 and st["hbm_in_use"] == 0
    assert st["iterations"] >= 1
    for k, (res, spin, inc, nu, s) in enumerate(MIX):
        d = np.abs(alone[k] - mixed[k])
        print(f"[serving] request {k}: alone vs mixed max abs diff {d.max():.3e} at {np.unravel_index(d.argmax(), d.shape)}")
        assert np.array_equal(alone[k], mixed[k]), k
        # same per-ray expressions as mk_render, with FMA contraction confined
        # to single expressions: the same bits
        ref = _render(models[s], res, spin, inc, nu, precision)
        assert s == 1 or np.count_nonzero(ref) > res
        d = np.abs(alone[k] - ref)
        print(f"[serving] {precision} request {k}: served vs mk_render max abs diff {d.max():.2e}")
        np.testing.assert_array_equal(alone[k], ref)
        if precision != "plain64" and s == 0:
            # the precision bar on the torus: flux within 1e-3 of plain FP64
            ref64 = _render(models[s], res, spin, inc, nu, "plain64")
            assert abs(alone[k].sum() - ref64.sum()) <= 1e-3 * ref64.sum()


@pytest.mark.parametrize("in_launch", [True, False])
@pytest.mark.parametrize("persistent", [False, True])
def test_served_shadow_equals_bisection_kernel(persistent, in_launch):
    """Bisections served inside the launches, one lane per angle
    (in_launch), and run by the idle server itself in the band schedule:
    mk_shadow_bisect's radii either way."""
    ang = np.linspace(0, 2 * np.pi, 40, endpoint=False)
    ref_a = find_shadow_bisection_angles(0.9, 60.0, ang, max_steps=2000)
    ref_b = find_shadow_bisection_angles(-0.4, 20.0, ang[::3], max_steps=1500, error_allowed=1e-2)
    with Server("mixed", lanes=1024, steps_per_launch=128, persistent=persistent,
                shadow_in_launch=in_launch) as srv:
        a = srv.submit_shadow(0.9, 60.0, ang, max_steps=2000)
        b = srv.submit_shadow(-0.4, 20.0, ang[::3], max_steps=1500, error_allowed=1e-2)
        c = srv.submit_shadow(0.1, 45.0, ang[:5], error_allowed=20.0)  # no halving: 0.5
        np.testing.assert_array_equal(srv.result(a, 120), ref_a)
        np.testing.assert_array_equal(srv.result(b, 120), ref_b)
        np.testing.assert_array_equal(srv.result(c, 120), np.full(5, 0.5))
    # and the oracle's bisection to 1e-6 M
    np.testing.assert_allclose(ref_a, c_oracle.find_shadow_bisection_angles(0.9, 60.0, ang, 2000, 1e-3, 40),
                               rtol=0, atol=1e-6)
    assert np.all((ref_a > 3) & (ref_a < 7))


@pytest.mark.parametrize("persistent", [False, True])
def test_idle_server_runs_small_bisections_itself(models, persistent):
    """An idle server runs a bisection of up to 1024 angles itself, in the
    band schedule, without a launch; a bisection queued behind images, and
    one of more than 1024 angles, go through the launches.  Every radius is
    mk_shadow_bisect's, and a wait on a request returns only once it is done."""
    small = np.linspace(0, 2 * np.pi, 64, endpoint=False)
    large = np.linspace(0, 2 * np.pi, 1100, endpoint=False)
    ref_s = find_shadow_bisection_angles(0.9, 60.0, small)
    ref_l = find_shadow_bisection_angles(0.5, 30.0, large, max_steps=1000)
    kw = _img_kw(128, 0.5, 30.0, 230e9)
    ref_i = _render(models[0], kw["resolution"], kw["bhspin"], kw["inclination"], kw["observing_frequency"], "mixed")
    with Server("mixed", persistent=persistent, image_in_launch=True) as srv:
        sid = srv.add_snapshot(models[0])
        t0 = time.perf_counter()
        np.testing.assert_array_equal(srv.result(srv.submit_shadow(0.9, 60.0, small), timeout=60), ref_s)
        t_small = time.perf_counter() - t0
        st = srv.stats()
        assert st["iterations"] == 0 and st["completed"] == 1  # no launch
        img = srv.submit_image(sid, **kw)
        sh = srv.submit_shadow(0.9, 60.0, small)     # behind an image: through a launch
        big = srv.submit_shadow(0.5, 30.0, large, max_steps=1000)  # over 1024 angles: through a launch
        np.testing.assert_array_equal(srv.result(img, timeout=60), ref_i)
        np.testing.assert_array_equal(srv.result(sh, timeout=60), ref_s)
        np.testing.assert_array_equal(srv.result(big, timeout=60), ref_l)
        time.sleep(0.1)
        st = srv.stats()
    print(f"[serving] idle 64-angle bisection: {t_small * 1e3:.2f} ms; {st['iterations']} launches after")
    assert st["completed"] == 4 and st["in_flight"] == 0 and st["iterations"] >= 1


@pytest.mark.parametrize("persistent", [False, True])
def test_idle_server_renders_a_lone_small_image_itself(models, persistent):
    """An idle server whose queue holds one image of up to 512^2 renders it
    itself through mk_render, without a launch; two images queued together,
    and a larger image, go through the launches.  Every image is mk_render's,
    bit for bit."""
    kws = [_img_kw(256, 0.5, 30.0, 230e9), _img_kw(128, -0.6, 70.0, 86e9), _img_kw(512, 0.9, 60.0, 345e9),
           _img_kw(640, 0.3, 45.0, 230e9)]
    refs = [_render(models[0], kw["resolution"], kw["bhspin"], kw["inclination"], kw["observing_frequency"], "mixed")
            for kw in kws]
    with Server("mixed", persistent=persistent) as srv:
        sid = srv.add_snapshot(models[0])
        for kw, ref in zip(kws[:3], refs[:3]):
            t0 = time.perf_counter()
            np.testing.assert_array_equal(srv.result(srv.submit_image(sid, **kw), timeout=60), ref)
            print(f"[serving] lone {kw['resolution']}^2 image: {(time.perf_counter() - t0) * 1e3:.2f} ms")
        st = srv.stats()
        assert st["iterations"] == 0 and st["completed"] == 3
        pair = [srv.submit_image(sid, **kw) for kw in kws[:2]]   # queued together: a launch
        big = srv.submit_image(sid, **kws[3])                     # over 512^2: a launch
        for rid, ref in zip(pair + [big], refs[:2] + refs[3:]):
            np.testing.assert_array_equal(srv.result(rid, timeout=60), ref)
        time.sleep(0.1)
        st = srv.stats()
    assert st["completed"] == 6 and st["in_flight"] == 0 and st["iterations"] >= 1


@pytest.mark.parametrize("persistent", [False, True])
def test_priorities_and_cancellation(models, persistent):
    """With an output budget for one request at a time, queued requests are
    admitted by priority, then in submit order; a cancelled queued request is
    never run and a wait on it reports the cancellation; a request that is
    running or done cannot be cancelled.  Every served image is mk_render's."""
    res = 48
    ang = np.linspace(0, 2 * np.pi, 1024, endpoint=False)
    budget = 3 * ang.size * 8 + 1024         # the bisection's staging; one 48^2 image fits, two do not
    assert res * res * 8 + 256 <= budget < 2 * (res * res * 8 + 256)
    ref = _render(models[0], res, 0.5, 30.0, 230e9, "mixed")
    with Server("mixed", hbm_budget=budget, persistent=persistent, shadow_in_launch=True,
                image_in_launch=True) as srv:
        sid = srv.add_snapshot(models[0])
        a = srv.submit_shadow(0.9, 60.0, ang)                           # holds the budget for ~30 ms
        t_wait = time.perf_counter() + 30.0
        while srv.stats()["admitted"] < 1:                               # a is in flight
            assert time.perf_counter() < t_wait, "the bisection was not admitted"
            time.sleep(0.0005)
        kw = _img_kw(res, 0.5, 30.0, 230e9)
        b = srv.submit_image(sid, **kw)
        c = srv.submit_image(sid, **kw)
        d = srv.submit_image(sid, priority=5, **kw)
        e = srv.submit_image(sid, **kw)
        assert srv.cancel(c) is True
        for rid in (b, d, e):
            np.testing.assert_array_equal(srv.result(rid, timeout=120), ref)
        srv.result(a, timeout=120)
        with pytest.raises(_lib.MkError, match="cancelled"):
            srv.wait(c, timeout=1.0)
        assert srv.cancel(a) is False and srv.cancel(b) is False  # done
        with pytest.raises(_lib.MkError, match="unknown request"):
            srv.cancel(10 ** 6)
        order = [i for i, _ in srv.poll()]
        st = srv.stats()
    print(f"[serving] completion order {order} (a={a}, b={b}, c={c} cancelled, d={d} priority 5, e={e})")
    assert order == [a, d, b, e]
    assert st["submitted"] == 5 and st["completed"] == 4 and st["queued"] == 0 and st["in_flight"] == 0


@pytest.mark.parametrize("persistent", [False, True])
def test_random_requests_priorities_and_cancellations(models, persistent):
    """48 requests drawn at random (seed 7) from two submitting threads:
    images of 8-136^2 on either snapshot at random spin, inclination and
    frequency, and bisections of 1-300 angles, with random priorities, some
    cancelled right after submission.  Every request that is not cancelled
    (cancel() returned True) equals mk_render's image or mk_shadow_bisect's
    radii bit for bit, whichever path served it (a launch, or the idle
    worker), and the counts add up."""
    import threading
    rng = np.random.Generator(np.random.PCG64(7))
    specs = []
    for k in range(48):
        if rng.random() < 0.7:
            specs.append(("img", int(rng.integers(1, 18)) * 8, float(rng.uniform(-0.95, 0.95)),
                          float(rng.uniform(5, 85)), float(rng.choice([86e9, 230e9, 345e9])), int(rng.integers(0, 2)),
                          int(rng.integers(-2, 3)), bool(rng.random() < 0.2)))
        else:
            specs.append(("sh", int(rng.integers(1, 301)), float(rng.uniform(-0.95, 0.95)), float(rng.uniform(5, 85)),
                          0.0, 0, int(rng.integers(-2, 3)), bool(rng.random() < 0.2)))
    refs = {}
    for k, (kind, n, a, inc, nu, snap, prio, _) in enumerate(specs):
        if kind == "img":
            refs[k] = _render(models[snap], n, a, inc, nu, "mixed")
        else:
            refs[k] = find_shadow_bisection_angles(a, inc, np.arange(n) / n * 2 * np.pi, max_steps=1500)
    got, cancelled, errors = {}, set(), []
    with Server("mixed", persistent=persistent) as srv:
        sids = [srv.add_snapshot(m) for m in models]

        def client(ks):
            try:
                mine = []
                for k in ks:
                    kind, n, a, inc, nu, snap, prio, cancel = specs[k]
                    if kind == "img":
                        rid = srv.submit_image(sids[snap], priority=prio, **_img_kw(n, a, inc, nu))
                    else:
                        rid = srv.submit_shadow(a, inc, np.arange(n) / n * 2 * np.pi, max_steps=1500, priority=prio)
                    if cancel and srv.cancel(rid):
                        cancelled.add(k)
                    else:
                        mine.append((k, rid))
                for k, rid in mine:
                    got[k] = srv.result(rid, timeout=120)
            except Exception as e:  # reported below, on the main thread
                errors.append(repr(e))

        threads = [threading.Thread(target=client, args=(range(t, 48, 2),)) for t in range(2)]
        for th in threads:
            th.start()
        for th in threads:
            th.join(timeout=300)
        srv.wait(-1, timeout=120)
        st = srv.stats()
    assert not errors, errors
    print(f"[serving] random requests: {len(got)} served, {len(cancelled)} cancelled, {st['iterations']} launches")
    assert sorted(set(got) | cancelled) == list(range(48)) and not (set(got) & cancelled)
    for k, v in got.items():
        np.testing.assert_array_equal(v, refs[k], err_msg=f"request {k}: {specs[k]}")
    assert st["submitted"] == 48 and st["completed"] == len(got) and st["queued"] == 0 and st["in_flight"] == 0


def test_persistent_launch_ends_when_idle_and_resumes_its_tickets():
    """The persistent launch ends after the server has been idle; its waves
    leave holding unpublished tickets, and the next launch resumes at the
    published count, so requests after each relaunch are served in full."""
    ang = np.linspace(0, 2 * np.pi, 200, endpoint=False)  # 4 tiles, the last one partial
    ref = find_shadow_bisection_angles(0.7, 50.0, ang, max_steps=1500)
    with Server("mixed", lanes=4096, persistent=True, shadow_in_launch=True) as srv:
        for k in range(4):
            print(f"[serving] persistent relaunch round {k}", flush=True)
            rid = srv.submit_shadow(0.7, 50.0, ang, max_steps=1500)
            np.testing.assert_array_equal(srv.result(rid, 60), ref)
            time.sleep(0.1)  # longer than the idle limit: the launch ends
        st = srv.stats()
    assert st["completed"] == 4 and st["in_flight"] == 0
    assert st["iterations"] >= 4  # one launch per round


def test_persistent_bisection_arriving_during_an_image_only_launch(models):
    """Default lanes, mixed: a launch that serves only images runs three waves
    per SIMD without the bisection path.  A bisection submitted while it runs
    waits until that launch has drained; the next launch serves both kinds.
    Every result is the same as when served alone."""
    ang = np.linspace(0, 2 * np.pi, 100, endpoint=False)
    ref_sh = find_shadow_bisection_angles(0.6, 40.0, ang, max_steps=1500)
    kws = [_img_kw(256, 0.5, 30.0, 230e9), _img_kw(192, -0.3, 60.0, 86e9), _img_kw(128, 0.8, 20.0, 345e9)]
    refs = [_render(models[0], kw["resolution"], kw["bhspin"], kw["inclination"], kw["observing_frequency"], "mixed")
            for kw in kws]
    with Server("mixed", persistent=True, shadow_in_launch=True, image_in_launch=True) as srv:
        sid = srv.add_snapshot(models[0])
        first = [srv.submit_image(sid, **kw) for kw in kws]
        t_wait = time.perf_counter() + 30.0
        while srv.stats()["iterations"] < 1:  # the image-only launch has started
            assert time.perf_counter() < t_wait, "the first launch did not start"
            time.sleep(0.0005)
        sh = srv.submit_shadow(0.6, 40.0, ang, max_steps=1500)
        later = [srv.submit_image(sid, **kw) for kw in kws]
        for rid, ref in zip(first + later, refs + refs):
            np.testing.assert_array_equal(srv.result(rid, timeout=120), ref)
        np.testing.assert_array_equal(srv.result(sh, timeout=120), ref_sh)
        time.sleep(0.1)
        st = srv.stats()
    print(f"[serving] image-only then bisection: {st['iterations']} launches")
    assert st["completed"] == 7 and st["in_flight"] == 0
    assert st["iterations"] >= 2


@pytest.mark.parametrize("persistent", [False, True])
def test_concurrent_submitters(models, persistent):
    """Four host threads submit, wait and read statistics at once (the C
    calls release the GIL): every image equals mk_render's, and the counts
    add up."""
    import threading
    specs = [(32 + 16 * (k % 3), 0.5 - 0.2 * (k % 4), 20.0 + 15 * (k % 5), (86e9, 230e9, 345e9)[k % 3], k % 2)
             for k in range(20)]
    refs = {k: _render(models[s], res, spin, inc, nu, "mixed") for k, (res, spin, inc, nu, s) in enumerate(specs)}
    got, errors = {}, []
    with Server("mixed", persistent=persistent) as srv:
        sids = [srv.add_snapshot(m) for m in models]

        def client(ks):
            try:
                for k in ks:
                    res, spin, inc, nu, s = specs[k]
                    rid = srv.submit_image(sids[s], **_img_kw(res, spin, inc, nu))
                    srv.stats()
                    got[k] = srv.result(rid, timeout=120)
            except Exception as e:  # reported below, on the main thread
                errors.append(repr(e))

        threads = [threading.Thread(target=client, args=(range(t, 20, 4),)) for t in range(4)]
        for th in threads:
            th.start()
        for th in threads:
            th.join(timeout=300)
        st = srv.stats()
    assert not errors, errors
    assert sorted(got) == list(range(20))
    for k in range(20):
        np.testing.assert_array_equal(got[k], refs[k])
    assert st["submitted"] == 20 and st["completed"] == 20 and st["in_flight"] == 0


@pytest.mark.parametrize("persistent", [False, True])
def test_budget_back_pressure_and_errors(models, persistent):
    res = 48
    # one request: its output (256-byte aligned already) and its tile order
    # (36 tiles x 4 bytes, aligned to 256; the persistent form keeps the
    # order outside the budget)
    need = res * res * 8 + 256
    with Server("mixed", hbm_budget=need, lanes=4096, persistent=persistent, image_in_launch=True) as srv:
        sid = srv.add_snapshot(models[0])
        ref = _render(models[0], res, 0.5, 30.0, 230e9, "mixed")
        rids = [srv.submit_image(sid, **_img_kw(res, 0.5, 30.0, 230e9)) for _ in range(3)]
        st = srv.stats()
        assert st["hbm_budget"] == need and st["hbm_in_use"] <= need
        outs = [srv.result(r, timeout=120) for r in rids]
        for o in outs:  # served = mk_render, bit for bit, under back-pressure too
            np.testing.assert_array_equal(o, ref)
        with pytest.raises(_lib.MkError, match="budget"):
            srv.submit_image(sid, **_img_kw(64, 0.5, 30.0, 230e9))
        with pytest.raises(_lib.MkError, match="unknown snapshot"):
            srv.submit_image(7, **_img_kw(16, 0.5, 30.0, 230e9))
        done = srv.poll()
        assert sorted(i for i, _ in done) == sorted(rids)
        assert all(ms > 0 for _, ms in done)
        srv.wait(rids[0], timeout=0.0)  # finished and polled: returns at once
        with pytest.raises(_lib.MkError, match="unknown request"):
            srv.wait(10 ** 6, timeout=0.0)
    with Server("plain64", persistent=persistent) as srv:
        sid = srv.add_snapshot(models[0])
        rid = srv.submit_image(sid, **_img_kw(256, 0.5, 30.0, 230e9))
        with pytest.raises(_lib.MkError, match="timed out"):
            srv.wait(rid, timeout=0.0)
        srv.wait(-1, timeout=120)
        assert srv.stats()["completed"] == 1
