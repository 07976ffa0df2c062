#!/usr/bin/env python3
"""Benchmark of the per-ray hot path on one MI355X.

Workload: 256 x 256 image, a = 0.5, inclination 30 deg, 230 GHz, through a
synthetic analytic torus stored as 512 meshblocks of 32^3 FP32 cells over
+-40 M (a uniform 256^3 grid).  A step is one image: one launch of the fused
kernel (camera ray, RK4 geodesic, sampling, emission, transfer) over all
65 536 rays, inputs resident in HBM.  The timed precision is the mixed
FP32/FP64 kernel; the plain FP64 kernel is timed the same way beside it.

Prints one JSON line.  rays/s counts camera rays fully integrated.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

MSUN = 1.989e33
FLOP_PER_STEP = 1000.0        # algorithmic FLOP per accepted RK4 ray-step
PEAK_FP64 = 78.6              # TFLOP/s, MI355X FP64 vector (datasheet)
PEAK_FP32 = 157.3             # TFLOP/s, MI355X FP32 vector (datasheet)
HBM_PEAK = 8000.0             # GB/s


def ring_diameter(img, fov):
    """Twice the radius of the peak of the azimuthally averaged profile."""
    res = img.shape[0]
    c = np.linspace(-fov / 2, fov / 2, 2 * res + 1)[1::2]
    x, y = np.meshgrid(c, c, indexing="ij")
    r = np.hypot(x, y)
    bins = np.linspace(0, fov / 2, res // 2 + 1)
    idx = np.digitize(r.ravel(), bins) - 1
    ok = (idx >= 0) & (idx < res // 2)
    prof = np.bincount(idx[ok], img.ravel()[ok], res // 2) / np.maximum(np.bincount(idx[ok], None, res // 2), 1)
    k = int(np.argmax(prof))
    return float(bins[k] + bins[k + 1])


def time_launches(L, fn, steps, warmup):
    """Per-launch durations (ms) from HIP events on the launch stream, plus
    the wall time of the whole timed region bracketed by device syncs."""
    import ctypes as C
    for _ in range(warmup):
        fn()
    L.check(L.lib().mk_stream_sync(None))
    evs = []
    for _ in range(2 * steps):
        e = C.c_void_p()
        L.check(L.lib().mk_event_create(C.byref(e)))
        evs.append(e.value)
    t0 = time.perf_counter()
    for s in range(steps):
        L.check(L.lib().mk_event_record(evs[2 * s], None))
        fn()
        L.check(L.lib().mk_event_record(evs[2 * s + 1], None))
    L.check(L.lib().mk_stream_sync(None))
    wall = time.perf_counter() - t0
    ms = []
    for s in range(steps):
        v = C.c_float()
        L.check(L.lib().mk_event_elapsed_ms(evs[2 * s], evs[2 * s + 1], C.byref(v)))
        ms.append(v.value)
    for e in evs:
        L.lib().mk_event_destroy(e)
    return np.array(ms), wall


def serve_stream(models, precision, seconds, rate, seed=0, log=None, persistent=False):
    """Poisson stream of image requests against resident snapshots through
    the request server: resolution uniform in {64, 128, 256, 512}^2, spin
    U[-0.98, 0.98], inclination U[5, 85] deg, frequency in {86, 230, 345} GHz,
    snapshot uniform over the resident ones.  Returns throughput and
    submit-to-result latency percentiles."""
    import threading
    from mahakala_amd.serving import Server

    rng = np.random.Generator(np.random.PCG64(seed))
    n_req = max(1, int(round(seconds * rate)))
    t_arr = np.cumsum(rng.exponential(1.0 / rate, n_req))
    res = rng.choice([64, 128, 256, 512], n_req)
    spin = rng.uniform(-0.98, 0.98, n_req)
    inc = rng.uniform(5.0, 85.0, n_req)
    nu = rng.choice([86e9, 230e9, 345e9], n_req)
    snap = rng.integers(0, len(models), n_req)
    with Server(precision, use_graph=True, persistent=persistent) as srv:
        sids = [srv.add_snapshot(m) for m in models]
        # one small request first so that the first timed request does not
        # pay for lazy initialisation
        srv.result(srv.submit_image(sids[0], 0.5, 30.0, resolution=16), timeout=120)
        srv.poll()
        st0 = srv.stats()
        ids = []
        t0 = time.perf_counter()

        def submitter():
            for k in range(n_req):
                dt = t0 + t_arr[k] - time.perf_counter()
                if dt > 0:
                    time.sleep(dt)
                ids.append(srv.submit_image(sids[snap[k]], float(spin[k]), float(inc[k]), resolution=int(res[k]),
                                            observing_frequency=float(nu[k]), fluid_gamma=13. / 9))

        th = threading.Thread(target=submitter)
        th.start()
        th.join()
        srv.wait(-1, timeout=600)
        wall = time.perf_counter() - t0
        done = dict(srv.poll(cap=n_req + 16))
        if persistent:
            time.sleep(0.1)  # the idle launch ends and posts its counters
        st = srv.stats()
    lat = np.array([done[i] for i in ids])
    rays = float(np.sum(res.astype(np.float64) ** 2))
    out = {"workload": f"Poisson arrivals at {rate:g} req/s for {seconds:g} s ({n_req} image requests), "
                       "resolution uniform in {64,128,256,512}^2, spin U[-0.98,0.98], inclination U[5,85] deg, "
                       f"frequency in {{86,230,345}} GHz, {len(models)} resident 256^3-cell torus snapshots",
           "precision": precision, "form": "persistent launch" if persistent else "relaunched every 256 steps",
           "requests": n_req, "rays": int(rays), "wall_s": round(wall, 3),
           "rays_per_s": round(rays / wall, 1), "offered_rays_per_s": round(rays / t_arr[-1], 1),
           "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3), "p99": round(float(np.percentile(lat, 99)), 3),
                          "max": round(float(lat.max()), 3)},
           "gpu_busy_frac": round((st["gpu_ms"] - st0["gpu_ms"]) / 1e3 / wall, 4),
           "lane_utilisation": round((st["lane_steps_busy"] - st0["lane_steps_busy"])
                                     / max(1, st["lane_steps_total"] - st0["lane_steps_total"]), 4),
           "launches": int(st["iterations"] - st0["iterations"])}
    if log:
        print(f"[bench] serving {precision} ({out['form']}): {out['rays_per_s'] / 1e6:.2f} M rays/s, "
              f"p50 {out['latency_ms']['p50']} ms, "
              f"p99 {out['latency_ms']['p99']} ms, busy {out['gpu_busy_frac']}", file=log, flush=True)
    return out


def _heavy_snapshot(seed):
    """Turbulent torus snapshot number `seed` at its own spin, U[-0.98, 0.98]."""
    from mahakala_amd.grmhd.synthetic import torus_snapshot
    a = float(np.random.Generator(np.random.PCG64(1000 + seed)).uniform(-0.98, 0.98))
    return a, torus_snapshot(a, layout="uniform", seed=seed, turbulence_sigma=0.5)


def heavy_stream(models, seed=1, n_images=128, n_shadow=128, res=512, n_angles=4096, log=None, persistent=False):
    """256 queued requests against 8 resident snapshots: 512^2 images at 8
    frequencies log-spaced over 86-690 GHz (each image at its snapshot's spin
    and a random inclination) and 4096-angle shadow bisections at random spin
    and inclination, run once in mixed and once in plain FP64.  Reports
    throughput of both and the per-image flux error of mixed against plain
    FP64."""
    from mahakala_amd.serving import Server

    rng = np.random.Generator(np.random.PCG64(seed))
    freqs = np.geomspace(86e9, 690e9, 8)
    kinds = rng.permutation(np.r_[np.zeros(n_images, int), np.ones(n_shadow, int)])
    spin = rng.uniform(-0.98, 0.98, kinds.size)
    inc = rng.uniform(5.0, 85.0, kinds.size)
    nu = rng.choice(freqs, kinds.size)
    snap = rng.integers(0, len(models), kinds.size)
    angles = np.arange(n_angles) / n_angles * 2 * np.pi
    res_out, lat = {}, {}
    for precision in ("mixed", "plain64"):
        with Server(precision, use_graph=True, persistent=persistent) as srv:
            sids = [srv.add_snapshot(m) for m in models]
            srv.result(srv.submit_image(sids[0], 0.5, 30.0, resolution=16), timeout=120)
            srv.poll()
            st0 = srv.stats()
            t0 = time.perf_counter()
            ids = []
            for k in range(kinds.size):
                if kinds[k] == 0:
                    ids.append(srv.submit_image(sids[snap[k]], models[snap[k]].bhspin, float(inc[k]), resolution=res,
                                                observing_frequency=float(nu[k]), fluid_gamma=13. / 9))
                else:
                    ids.append(srv.submit_shadow(float(spin[k]), float(inc[k]), angles))
            srv.wait(-1, timeout=900)
            wall = time.perf_counter() - t0
            if persistent:
                time.sleep(0.1)  # the idle launch ends and posts its counters
            st = srv.stats()
            done = dict(srv.poll(cap=kinds.size + 16))
            res_out[precision] = [srv.result(i, 1) for i in ids]
            lat[precision] = dict(wall=wall, busy=(st["lane_steps_busy"] - st0["lane_steps_busy"])
                                  / max(1, st["lane_steps_total"] - st0["lane_steps_total"]),
                                  gpu=(st["gpu_ms"] - st0["gpu_ms"]) / 1e3,
                                  lat=np.array([done[i] for i in ids]))
        if log:
            print(f"[bench] heavy {precision}: {lat[precision]['wall']:.2f} s", file=log, flush=True)
    img = np.flatnonzero(kinds == 0)
    sh = np.flatnonzero(kinds == 1)
    sums = {p: np.array([float(res_out[p][k].sum()) for k in img]) for p in res_out}
    with np.errstate(all="ignore"):
        flux_err = np.abs(sums["mixed"] - sums["plain64"]) / np.abs(sums["plain64"])
    # The reference's explicit-Euler transfer, I += -dt L (j - alpha I), is
    # unstable where alpha |dt| L > 2 (optically thick steps, here the
    # low-frequency images of the densest turbulent snapshots): intensities
    # then turn negative and grow without bound in either precision, and the
    # two precisions diverge.  Such images are counted apart.
    stable = np.array([np.all(res_out["plain64"][k] >= 0) and np.all(np.isfinite(res_out["plain64"][k]))
                       for k in img])
    bad = [dict(spin=round(float(models[snap[k]].bhspin), 4), inc=round(float(inc[k]), 2), nu=float(nu[k]), snapshot=int(snap[k]),
                flux_mixed=float(sums["mixed"][n]), flux_plain=float(sums["plain64"][n]),
                nonfinite_mixed=int(np.count_nonzero(~np.isfinite(res_out["mixed"][k]))),
                nonfinite_plain=int(np.count_nonzero(~np.isfinite(res_out["plain64"][k]))),
                max_mixed=float(np.nanmax(res_out["mixed"][k])), max_plain=float(np.nanmax(res_out["plain64"][k])))
           for n, k in enumerate(img) if stable[n] and not flux_err[n] <= 1e-3]
    same_radii = all(np.array_equal(res_out["mixed"][k], res_out["plain64"][k]) for k in sh)
    iters = 14  # bracket [0.5, 10] halved to 1e-3
    out = {"workload": f"{kinds.size} queued requests against {len(models)} resident 512-block torus snapshots "
                       f"with log-normal turbulence (sigma 0.5, seeds 0-{len(models) - 1}, spins "
                       f"{[round(m.bhspin, 3) for m in models]}): {n_images} images of {res}^2 at 8 frequencies "
                       f"86-690 GHz at their snapshot's spin, and {n_shadow} shadow bisections of {n_angles} angles "
                       "at spin U[-0.98,0.98]; inclination U[5,85] deg",
           "image_rays": int(img.size * res * res), "bisection_ray_integrations": int(sh.size * n_angles * iters)}
    for precision in ("mixed", "plain64"):
        L = lat[precision]
        out[precision] = {"wall_s": round(L["wall"], 3), "gpu_busy_s": round(L["gpu"], 3),
                          "lane_utilisation": round(L["busy"], 4),
                          "requests_per_s": round(kinds.size / L["wall"], 2),
                          "latency_ms": {"p50": round(float(np.percentile(L["lat"], 50)), 2),
                                         "p99": round(float(np.percentile(L["lat"], 99)), 2)}}
    out["speedup_mixed_over_plain"] = round(lat["plain64"]["wall"] / lat["mixed"]["wall"], 3)
    out["image_flux_rel_err_mixed_vs_plain"] = {
        "stable_images": int(stable.sum()), "max": float(np.max(flux_err[stable])) if stable.any() else None,
        "median": float(np.median(flux_err[stable])) if stable.any() else None, "stable_images_above_1e-3": bad,
        "unstable_images": int((~stable).sum()),
        "unstable_median": float(np.median(flux_err[~stable])) if (~stable).any() else None}
    out["shadow_radii_identical"] = bool(same_radii)
    return out


def stage_stream(L, res=256, reps=3):
    """The unfused path's standalone kernels on the config-2 sample stream:
    the materialised history of the 256^2 image (T rows of every ray, as
    geodesic_integrator returns it), the sampler k_sample_scalars over all
    T x npx samples (stage b, athenak.py:639-812), then k_synchrotron over the
    same samples and k_solve over the (T, npx) arrays (stage c,
    transfer.py:30-119), each launched `reps` times and timed with HIP events.
    Reports each kernel's time and its algorithmic bytes per second against
    8 TB/s: sampler 136 B per sample (state 64 B in, one 32-byte cell record,
    5 scalars 40 B out), synchrotron 56 B (5 inputs, j and alpha), solve 24 B
    per sample (j, alpha, dt) + 8 B per ray."""
    import ctypes as C
    from mahakala_amd.grmhd.athenak import AthenakFluidModel
    from mahakala_amd.grmhd.synthetic import uniform_torus_snapshot
    a, inc, nu, fov = 0.5, 30.0, 230e9, 20.0
    model = AthenakFluidModel.from_datasets(uniform_torus_snapshot(a, seed=0), a, fluid_gamma=13. / 9)
    lib = L.lib()
    n = res * res
    d0 = L.DeviceBuffer(8 * 8 * n)
    L.check(lib.mk_camera_grid(a, inc, 1000.0, -fov / 2, fov / 2, res, d0.ptr, None))
    dK = L.DeviceBuffer(4 * n)
    L.check(lib.mk_integrate_final(d0.ptr, n, 10000, 40.0, 1e-4, a, dK.ptr, None, None, None))
    T = C.c_long()
    L.check(lib.mk_trim_rows(dK.ptr, n, 10000, C.byref(T), None))
    T = T.value
    ns = T * n
    dS, dD = L.DeviceBuffer(8 * 8 * ns), L.DeviceBuffer(8 * ns)
    L.check(lib.mk_integrate_history(d0.ptr, n, 10000, 40.0, 1e-4, a, T, dS.ptr, dD.ptr, None))
    dsc = L.DeviceBuffer(5 * 8 * ns)

    def sampler():
        L.check(lib.mk_sample_scalars(model.device_snapshot.handle, dS.ptr, ns, a, np.pi / 3, dsc.ptr, None))
    ms_b, _ = time_launches(L, sampler, reps, 1)
    # the per-sample glue of make_image (images.py:86-118) on the host, as the
    # unfused path does, to give the emission kernel its real inputs
    sc = dsc.to_array((5, ns))
    del dS
    units = model.get_units(6.2e9 * MSUN, 1e26)
    with np.errstate(all="ignore"):
        bsq = sc[4] * sc[4]
        beta = sc[1] * (13. / 9 - 1.) / bsq / 0.5
        tr = (40.0 * beta * beta + 1.0) / (1.0 + beta * beta)
        theta = (2.99792458e10 ** 2 * (1.6726e-24 * sc[1] * (1 / 3.) * (2 / 3.))
                 / (sc[0] * (2 / 3. + (1 / 3.) * tr)) / (9.1094e-28 * 2.99792458e10 ** 2))
        ins = [units["Ne_unit"] * sc[0], theta, units["B_unit"] * sc[4], sc[2], -sc[3] * nu]
    del sc, bsq, beta, tr, theta
    dins = [L.DeviceBuffer.from_array(v) for v in ins]
    del ins
    dj, dal = L.DeviceBuffer(8 * ns), L.DeviceBuffer(8 * ns)
    dI = L.DeviceBuffer(8 * n)

    def synch():
        L.check(lib.mk_synchrotron(*[b.ptr for b in dins], ns, 1, 1.0 / nu, dj.ptr, dal.ptr, None))

    def solve():
        L.check(lib.mk_solve_intensity(dj.ptr, dal.ptr, dD.ptr, T, n, units["L_unit"], dI.ptr, None, None))
    ms_s, _ = time_launches(L, synch, reps, 1)
    ms_t, _ = time_launches(L, solve, reps, 1)
    out = {"workload": f"materialised history of the {res}^2 config-2 image (a=0.5, i=30 deg, 230 GHz, torus on "
                       f"512 blocks of 32^3): T = {T} rows x {n} rays = {ns} samples",
           "samples": int(ns), "T": int(T)}
    for name, ms, bps, extra in (("k_sample_scalars", ms_b, 136, 0), ("k_synchrotron", ms_s, 56, 0),
                                 ("k_solve", ms_t, 24, 8 * n)):
        t = float(np.mean(ms))
        gbs = (bps * ns + extra) / (t * 1e-3) / 1e9
        out[name] = {"ms": round(t, 4), "algorithmic_bytes": int(bps * ns + extra),
                     "algorithmic_GBps": round(gbs, 1), "frac_of_8TBps": round(gbs / HBM_PEAK, 4)}
        print(f"[stages] {name}: {t:.3f} ms, {gbs:.0f} GB/s algorithmic", file=sys.stderr, flush=True)
    return out


def step_cycles(L, steps=2000, repeats=3):
    """One wave alone on a SIMD repeats each stage of the per-ray loop
    (mk_step_cycles): shader cycles per repetition, FP64 against mixed, and
    the mixed-over-FP64 ratios they allow -- for the chain of one ray (RK4 +
    radius + step rule: what bounds a launch whose longest ray sets its time)
    and for a whole accepted step (geodesic + interpolation + emission: what
    bounds a launch that keeps every SIMD busy)."""
    names = ["rk4", "rk4", "geodesic_step", "geodesic_step", "interp", "interp", "interp_emission",
             "interp_emission"]
    out = {"fp64": {}, "mixed": {}}
    ghz = []
    for v in range(8):
        best = None
        for _ in range(repeats):
            h = np.zeros(3)
            L.check(L.lib().mk_step_cycles(v, steps, 0.5, h.ctypes.data))
            best = h if best is None or h[0] < best[0] else best
        out["fp64" if v % 2 == 0 else "mixed"][names[v]] = round(float(best[0]), 1)
        ghz.append(float(best[2]))
    f, m = out["fp64"], out["mixed"]
    whole64 = f["geodesic_step"] + f["interp_emission"]
    wholem = m["geodesic_step"] + m["interp_emission"]
    out["what"] = ("shader cycles per repetition of one stage by one wave alone on a SIMD (2000 repetitions, best "
                   "of 3); interp_emission includes the interpolation")
    out["clock_ghz"] = round(float(np.median(ghz)), 3)
    out["ratio_fp64_over_mixed"] = {
        "rk4": round(f["rk4"] / m["rk4"], 3),
        "geodesic_chain": round(f["geodesic_step"] / m["geodesic_step"], 3),
        "interp_emission": round(f["interp_emission"] / m["interp_emission"], 3),
        "whole_step": round(whole64 / wholem, 3)}
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--res", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--only", choices=["mixed", "plain64", "mixed-zone"], default=None,
                    help="time one precision only (profiling runs)")
    ap.add_argument("--serve-seconds", type=float, default=5.0, help="length of the Poisson serving stream")
    ap.add_argument("--serve-rate", type=float, default=200.0, help="mean requests/s of the serving stream")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving stream")
    ap.add_argument("--serve-form", choices=["relaunch", "persistent", "both"], default="persistent",
                    help="server form for the Poisson and heavy streams (both: an A/B of the two)")
    ap.add_argument("--no-heavy", action="store_true", help="skip the heavy queued stream (8 snapshots)")
    ap.add_argument("--no-target", action="store_true", help="skip the a=0.9 precision-target image")
    ap.add_argument("--no-sizes", action="store_true", help="skip the 512^2 and 1024^2 images")
    ap.add_argument("--no-octree", action="store_true",
                    help="skip the 1024^2 image on the 3-level turbulent octree (a=0.94, i=17 deg)")
    ap.add_argument("--no-shadow", action="store_true", help="skip the shadow bisection timings")
    ap.add_argument("--form", choices=["auto", "classic", "paired", "triple"], default="auto",
                    help="render kernel form (auto: the library's choice)")
    ap.add_argument("--stages", action="store_true",
                    help="only time the unfused sampler, emission and transfer kernels on the config-2 sample stream")
    ap.add_argument("--step-cycles", action="store_true",
                    help="only time the per-step stages in one wave (FP64 against mixed) and print them")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run the device copy kernel over MIB MiB 3 times (traffic counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported (run with --gpus 1)",
              file=sys.stderr)
        sys.exit(2)

    from mahakala_amd import _lib as L
    from mahakala_amd import images
    if args.copy_calibration:
        nb = args.copy_calibration << 20
        src, dst = L.DeviceBuffer(nb), L.DeviceBuffer(nb)
        L.check(L.lib().mk_memset(src.ptr, 1, nb, None))
        for _ in range(3):
            L.check(L.lib().mk_device_copy(dst.ptr, src.ptr, nb, None))
        L.sync()
        print(json.dumps({"copy_calibration_bytes": nb, "launches": 3}))
        return
    if args.step_cycles:
        print(json.dumps(step_cycles(L)))
        return
    if args.stages:
        print(json.dumps(stage_stream(L)))
        return
    from mahakala_amd.grmhd.athenak import AthenakFluidModel
    from mahakala_amd.grmhd.synthetic import uniform_torus_snapshot

    a, inc, nu, fov, res = 0.5, 30.0, 230e9, 20.0, args.res
    t0 = time.perf_counter()
    ds = uniform_torus_snapshot(a, seed=0)
    model = AthenakFluidModel.from_datasets(ds, a, fluid_gamma=13. / 9)
    snap = model.device_snapshot
    print(f"[bench] snapshot ready in {time.perf_counter() - t0:.1f}s ({snap.nbytes / 2**20:.0f} MiB on device)",
          file=sys.stderr, flush=True)
    n = res * res
    img = L.DeviceBuffer(8 * n)
    stp = L.DeviceBuffer(4 * n)
    precisions = ("plain64", "mixed", "mixed-zone")
    form = None if args.form == "auto" else args.form
    par = {p: images.render_params(a, camera_inclination=inc, fov=fov, resolution=res, observing_frequency=nu,
                                   precision=p, form=form) for p in precisions}

    def kernel_name(p):
        import ctypes as C
        form = C.c_int32()
        L.check(L.lib().mk_render_form(snap.handle, par[p], C.byref(form)))
        m = "true" if par[p].precision == 1 else "false"
        if form.value == images.FORMS["triple"]:
            # the workload's 512-block tables are staged in LDS; ring B holds
            # 3 rounds in mixed, 2 in FP64; the workload's 1024 tiles make a
            # one-wave launch, which runs the zone-free build (0) unless the
            # preset has a zone (1)
            zone = 1 if m == "true" and (par[p].fp64_radius > 0 or par[p].ring_radius > 0) else 0
            return f"k_render_tri<{m}, true, {3 if m == 'true' else 2}, {zone}>"
        if form.value == images.FORMS["paired"]:
            return f"k_render_pair<{m}>"
        # classic: tables in LDS; mixed runs three workgroups per CU with in-place samples
        return f"k_render<{m}, true, {m}>"

    def launch(p, steps_out=False):
        def fn():
            L.check(L.lib().mk_render(snap.handle, par[p], img.ptr, stp.ptr if steps_out else None, None))
        return fn

    # reference run for accepted steps and accuracy (untimed)
    launch("plain64", True)()
    L.sync()
    I_plain = img.to_array((res, res))
    steps_plain = stp.to_array((res, res), np.int32)
    I_mix = {}
    for p in ("mixed", "mixed-zone"):
        launch(p)()
        L.sync()
        I_mix[p] = img.to_array((res, res))
    I_mixed = I_mix["mixed"]
    total_steps = float(steps_plain.sum())

    results = {}
    for p in precisions:
        if args.only and p != args.only:
            continue
        ms, wall = time_launches(L, launch(p), args.steps, args.warmup)
        results[p] = dict(ms=ms, wall=wall)
        print(f"[bench] {p}: {np.mean(ms):.3f} ms/image (p50 {np.percentile(ms, 50):.3f}, "
              f"p99 {np.percentile(ms, 99):.3f}), wall {wall:.3f}s for {args.steps}", file=sys.stderr, flush=True)

    flux_p = float(I_plain.sum())

    def accuracy(I):
        bright = I_plain > 1e-3 * I_plain.max()
        return dict(flux_rel_err=abs(float(I.sum()) - flux_p) / flux_p,
                    pixel_l1_rel=float(np.abs(I - I_plain).sum() / I_plain.sum()),
                    pixel_max_rel_bright=float((np.abs(I - I_plain)[bright] / I_plain[bright]).max()),
                    ring_diameter_plain=ring_diameter(I_plain, fov), ring_diameter_mixed=ring_diameter(I, fov))
    acc = accuracy(I_mixed)

    head = args.only or "mixed"
    ms_head = results[head]["ms"]
    wall = results[head]["wall"]
    value = n * args.steps / wall
    kern_ms = float(np.mean(ms_head))
    achieved = FLOP_PER_STEP * total_steps / (kern_ms * 1e-3) / 1e12
    peak = PEAK_FP32 if head == "mixed" else PEAK_FP64
    roofline = dict(bound="valu", achieved=round(achieved, 3), peak=peak, unit="TFLOP/s",
                    frac=round(achieved / peak, 4), traffic=None,
                    kernel=kernel_name(head),
                    algorithmic=f"{FLOP_PER_STEP:.0f} FLOP x {total_steps:.0f} accepted ray-steps per launch")
    tfile = os.path.join(ROOT, "profiles", "render_traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            tr = json.load(f).get(head)
        if tr and tr.get("kernel") == roofline["kernel"]:
            roofline["traffic"] = tr["bytes_per_launch"]
    # counter-derived VALU FLOP rate: SQ_INSTS_VALU_FLOPS_FP32 + _FP64 per
    # launch (wave-level, FMA counted twice) x 64 lanes over this run's kernel
    # time, from a PMC run of this same workload
    ffile = os.path.join(ROOT, "profiles", "render_flops.json")
    if os.path.exists(ffile):
        with open(ffile) as f:
            fl = json.load(f).get(head)
        if fl and fl.get("kernel") == roofline["kernel"]:
            flops = 64.0 * (fl["SQ_INSTS_VALU_FLOPS_FP32"] + fl["SQ_INSTS_VALU_FLOPS_FP64"])
            ach = flops / (kern_ms * 1e-3) / 1e12
            roofline["counters"] = {
                "flop_per_launch": flops, "achieved": round(ach, 3), "frac": round(ach / peak, 4),
                "frac_of_unpacked_peak": round(ach / (PEAK_FP32 / 2 if head != "plain64" else PEAK_FP64), 4),
                "what": "SQ_INSTS_VALU_FLOPS_FP32 + SQ_INSTS_VALU_FLOPS_FP64 (x 64 lanes) per launch from "
                        "profiles/render_flops.json over this run's HIP-event kernel time; the unpacked peak is "
                        "the FP32 rate without packed FP32 (78.6 TFLOP/s, as FP64)"}

    out = {
        "metric": "rays/s (256^2 Kerr + RT), mixed vs plain FP64 at 1e-3 flux err; p50/p99 latency",
        "value": round(value, 1), "unit": "rays/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * wall / args.steps, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32+fp64" if head == "mixed" else "fp64", "data": "synthetic",
        "config": {"workload": f"{res}x{res} image, a=0.5, i=30deg, 230 GHz, analytic torus on 512 meshblocks of 32^3 "
                               "FP32 cells over +-40 M",
                   "resolution": res, "rays_per_step": n, "precision": head, "max_nsteps": 10000},
        "image_launch_ms": {"p50": round(float(np.percentile(ms_head, 50)), 4),
                            "p99": round(float(np.percentile(ms_head, 99)), 4), "what": f"one {res}^2 image per launch"},
        "roofline": roofline,
        "accuracy_mixed_vs_plain": {k: (round(v, 9) if isinstance(v, float) else v) for k, v in acc.items()},
        "accepted_steps": int(total_steps), "mean_steps_per_ray": round(total_steps / n, 2),
        "max_steps_per_ray": int(steps_plain.max()),
    }
    if "plain64" in results:
        pw = results["plain64"]["wall"]
        pms = float(np.mean(results["plain64"]["ms"]))
        a64 = FLOP_PER_STEP * total_steps / (pms * 1e-3) / 1e12
        out["plain_fp64"] = {"value": round(n * args.steps / pw, 1), "unit": "rays/s", "ms_per_image": round(pms, 4),
                             "achieved_tflops": round(a64, 3), "peak": PEAK_FP64, "frac": round(a64 / PEAK_FP64, 4)}
        if "mixed" in results:
            out["speedup_mixed_over_plain"] = round(pw / results["mixed"]["wall"], 3)
        if "mixed-zone" in results:
            zw = results["mixed-zone"]["wall"]
            out["mixed_zone"] = {"what": "FP64 RK4 steps for an 8x8 tile's rays while any of them is inside r < 4 M or near the photon ring",
                                 "value": round(n * args.steps / zw, 1), "unit": "rays/s",
                                 "ms_per_image": round(float(np.mean(results["mixed-zone"]["ms"])), 4),
                                 "speedup_over_plain": round(pw / zw, 3),
                                 "accuracy_vs_plain": {k: (round(v, 9) if isinstance(v, float) else v)
                                                       for k, v in accuracy(I_mix["mixed-zone"]).items()}}

    if not args.no_sizes:
        # the same image at larger resolutions: more tiles than SIMDs, so the
        # launch is no longer bound by its longest ray
        sizes = {}
        for r2 in (512, 1024):
            img2 = L.DeviceBuffer(8 * r2 * r2)
            p2 = images.render_params(a, camera_inclination=inc, fov=fov, resolution=r2, observing_frequency=nu,
                                      precision=head if head != "mixed-zone" else "mixed")

            def fn2(p2=p2, img2=img2):
                L.check(L.lib().mk_render(snap.handle, p2, img2.ptr, None, None))
            ms2, w2 = time_launches(L, fn2, max(3, args.steps // 5), 2)
            sizes[f"{r2}x{r2}"] = {"ms_per_image": round(float(np.mean(ms2)), 3),
                                   "rays_per_s": round(r2 * r2 / (float(np.mean(ms2)) * 1e-3), 1)}
            del img2
        out["larger_images"] = dict(sizes, what="same scene and precision as value, one launch per image")
        print(f"[bench] larger images: {sizes}", file=sys.stderr, flush=True)

    if not args.no_octree:
        # 1024^2 at a = 0.94, i = 17 deg, 230 GHz through the 512-leaf 3-level
        # octree with log-normal turbulence: each precision timed, and the
        # mixed presets' error against plain FP64
        from mahakala_amd.grmhd.synthetic import octree_turbulent_snapshot
        t1 = time.perf_counter()
        mo = AthenakFluidModel.from_datasets(octree_turbulent_snapshot(0.94, seed=0), 0.94, fluid_gamma=13. / 9)
        print(f"[bench] octree snapshot ready in {time.perf_counter() - t1:.1f}s", file=sys.stderr, flush=True)
        ro = 1024
        imgo = L.DeviceBuffer(8 * ro * ro)
        oc, Io = {}, {}
        for pr in ("plain64", "mixed", "mixed-zone"):
            po = images.render_params(0.94, camera_inclination=17.0, resolution=ro, observing_frequency=230e9,
                                      precision=pr)

            def fno(po=po):
                L.check(L.lib().mk_render(mo.device_snapshot.handle, po, imgo.ptr, None, None))
            mso, _ = time_launches(L, fno, max(3, args.steps // 10), 2)
            Io[pr] = imgo.to_array((ro, ro))
            oc[pr] = {"ms_per_image": round(float(np.mean(mso)), 3),
                      "rays_per_s": round(ro * ro / (float(np.mean(mso)) * 1e-3), 1)}
        fo = float(Io["plain64"].sum())
        bro = Io["plain64"] > 1e-3 * Io["plain64"].max()
        for pr in ("mixed", "mixed-zone"):
            d = np.abs(Io[pr] - Io["plain64"])
            oc[pr].update(speedup_over_plain=round(oc["plain64"]["ms_per_image"] / oc[pr]["ms_per_image"], 3),
                          flux_rel_err=abs(float(Io[pr].sum()) - fo) / fo,
                          pixel_l1_rel=float(d.sum() / Io["plain64"].sum()),
                          pixel_max_rel_bright=float((d[bro] / Io["plain64"][bro]).max()),
                          ring_diameter=ring_diameter(Io[pr], 20.0))
        oc["plain64"]["ring_diameter"] = ring_diameter(Io["plain64"], 20.0)
        out["octree_1024"] = dict(oc, what="1024^2, a=0.94, i=17 deg, 230 GHz, fov 20 M, 512-leaf 3-level octree "
                                           "over +-50 M with log-normal turbulence (sigma 0.5, seed 0)")
        print(f"[bench] octree 1024^2: " + ", ".join(f"{k} {v['ms_per_image']} ms" for k, v in oc.items()),
              file=sys.stderr, flush=True)
        if not args.no_cpu:
            # config 3 on the host cores at a reduced resolution: the same view
            # at 256^2 through the C FP64 restatement, against the GPU's plain
            # FP64 image of that size
            from tests.helpers import c_snapshot_from_model
            rc3 = 256
            nth3 = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
            cso = c_snapshot_from_model(mo)
            t1 = time.perf_counter()
            I3, _ = cso.render(0.94, 17.0, 1000.0, 20.0, rc3, 10000, 13. / 9, 40.0, 1e26, 6.2e9 * MSUN, 230e9,
                               nthreads=nth3)
            dt3 = time.perf_counter() - t1
            p3 = images.render_params(0.94, camera_inclination=17.0, resolution=rc3, observing_frequency=230e9,
                                      precision="plain64")
            g3 = L.DeviceBuffer(8 * rc3 * rc3)
            L.check(L.lib().mk_render(mo.device_snapshot.handle, p3, g3.ptr, None, None))
            G3 = g3.to_array((rc3, rc3))
            f3 = float(G3.sum())
            out["cpu_baseline_config3"] = {
                "value": round(rc3 * rc3 / dt3, 1), "unit": "rays/s", "cores": nth3, "kind": "port",
                "sample": f"config 3's view at {rc3}x{rc3} instead of 1024x1024 (same camera, fov and octree), C FP64 "
                          f"restatement with OpenMP ({dt3:.1f} s wall)",
                "flux_rel_err_vs_gpu_plain": abs(float(I3.sum()) - f3) / f3}
            print(f"[bench] config 3 on the CPU at {rc3}^2: {rc3 * rc3 / dt3:.0f} rays/s", file=sys.stderr, flush=True)
            del cso, g3
        del mo, imgo

    if not args.no_shadow:
        # config 1's grid: 64x64 camera rays at a = 0.9, i = 60 deg integrated
        # with the reference's step rule (N = 10000, div 40, tol = 0.01 r_EH)
        # to their final accepted state, then the termination radius per pixel
        from mahakala_amd import geodesics as G
        a1 = 0.9
        tol1 = 0.01 * (1 + np.sqrt(1 - a1 * a1))
        s1 = np.asarray(G.initialize_geodesics_at_camera(a1, 60, 1000, -10, 10, 64))
        n1 = s1.shape[0]
        d1 = L.DeviceBuffer.from_array(s1)
        dK1, dl1, dr1 = L.DeviceBuffer(4 * n1), L.DeviceBuffer(64 * n1), L.DeviceBuffer(8 * n1)

        def fn1():
            L.check(L.lib().mk_integrate_final(d1.ptr, n1, 10000, 40.0, tol1, a1, dK1.ptr, dl1.ptr, None, None))
            L.check(L.lib().mk_radius_cal(dl1.ptr, n1, 8, a1, dr1.ptr, None))
        ms1, _ = time_launches(L, fn1, 10, 2)
        K1 = dK1.to_array((n1,), np.int32)
        out["config1_grid"] = {"ms": round(float(np.mean(ms1)), 3), "rays_per_s": round(n1 / (float(np.mean(ms1)) * 1e-3), 1),
                               "accepted_steps": int(np.minimum(K1, 10000).sum()), "max_steps_per_ray": int(K1.max()),
                               "what": "64x64 camera rays, a=0.9, i=60 deg, camera at 1000 M, N=10000, div 40, "
                                       "tol 0.01 r_EH: final accepted state and termination radius per pixel (FP64)"}
        print(f"[bench] config 1 grid 64^2: {np.mean(ms1):.3f} ms", file=sys.stderr, flush=True)
        del d1, dK1, dl1, dr1
        # find_shadow_bisection(0.9, 60, 64) and a 4096-angle job: 14 halvings,
        # each a full integration, reported as ray-integrations/s
        from mahakala_amd.geodesics import find_shadow_bisection_angles
        sh = {}
        for n_ang in (64, 4096):
            ang = np.arange(n_ang) / n_ang * 2 * np.pi
            find_shadow_bisection_angles(0.9, 60, ang)
            ts = []
            for _ in range(5):
                t0 = time.perf_counter()
                find_shadow_bisection_angles(0.9, 60, ang)
                ts.append(time.perf_counter() - t0)
            t = float(np.median(ts))
            sh[f"{n_ang}_angles"] = {"ms": round(t * 1e3, 3), "ray_integrations_per_s": round(n_ang * 14 / t, 1)}
        out["shadow_bisection"] = dict(sh, what="a=0.9, i=60 deg, N=2000, error 1e-3 (14 halvings), host arrays in "
                                                "and out (the reference call's boundary)")
        print(f"[bench] shadow bisection: {sh}", file=sys.stderr, flush=True)

    if not args.no_target:
        # the precision target's case: a = 0.9, 256^2, i = 60 deg (make_image's
        # default camera), torus generated at a = 0.9 on the 512-block layout
        ds9 = uniform_torus_snapshot(0.9, seed=0)
        m9 = AthenakFluidModel.from_datasets(ds9, 0.9, fluid_gamma=13. / 9)
        del ds9
        img9 = L.DeviceBuffer(8 * n)
        tgt = {}
        for p in ("plain64", "mixed", "mixed-zone"):
            rp9 = images.render_params(0.9, camera_inclination=60, fov=fov, resolution=res, observing_frequency=nu,
                                       precision=p)

            def fn9(rp9=rp9):
                L.check(L.lib().mk_render(m9.device_snapshot.handle, rp9, img9.ptr, None, None))
            fn9()
            L.sync()
            I9 = img9.to_array((res, res))
            ms9, w9 = time_launches(L, fn9, args.steps, args.warmup)
            tgt[p] = dict(ms=float(np.mean(ms9)), I=I9)
        f9 = float(tgt["plain64"]["I"].sum())
        out["target_a0.9"] = {
            "what": "256^2, a=0.9, i=60 deg, 230 GHz, torus at a=0.9 on 512 blocks of 32^3 over +-40 M",
            "plain64_ms": round(tgt["plain64"]["ms"], 4),
            **{p: {"ms": round(tgt[p]["ms"], 4), "speedup_over_plain": round(tgt["plain64"]["ms"] / tgt[p]["ms"], 3),
                   "flux_rel_err": abs(float(tgt[p]["I"].sum()) - f9) / f9} for p in ("mixed", "mixed-zone")}}
        print(f"[bench] a=0.9 target: plain {tgt['plain64']['ms']:.3f} ms, mixed {tgt['mixed']['ms']:.3f} ms",
              file=sys.stderr, flush=True)
        del m9, img9

    if not args.no_serve:
        ds1 = uniform_torus_snapshot(a, seed=1)
        model1 = AthenakFluidModel.from_datasets(ds1, a, fluid_gamma=13. / 9)
        forms = {"relaunch": [False], "persistent": [True], "both": [False, True]}[args.serve_form]
        for pers in forms:
            sv = serve_stream([model, model1], head if head != "mixed-zone" else "mixed", args.serve_seconds,
                              args.serve_rate, log=sys.stderr, persistent=pers)
            out["serving_persistent" if pers and len(forms) > 1 else "serving"] = sv
        sv = out["serving"]
        out["latency_ms"] = dict(sv["latency_ms"], what="submit-to-result latency of the serving stream")
        # one request at a time through an idle server (default flags, and with
        # every request kept inside the launches): submit-to-result medians
        from mahakala_amd.serving import Server
        ang64 = np.arange(64) / 64 * 2 * np.pi
        idle = {}
        for name, kw in (("default", {}), ("in_launch", {"image_in_launch": True, "shadow_in_launch": True})):
            with Server(head if head != "mixed-zone" else "mixed", **kw) as srv:
                sid = srv.add_snapshot(model)
                t_img, t_sh = [], []
                for _ in range(6):
                    time.sleep(0.05)  # longer than the idle limit: no launch running
                    t1 = time.perf_counter()
                    srv.result(srv.submit_image(sid, a, inc, resolution=256, observing_frequency=nu), timeout=60)
                    t_img.append(time.perf_counter() - t1)
                    time.sleep(0.05)
                    t1 = time.perf_counter()
                    srv.result(srv.submit_shadow(0.9, 60.0, ang64), timeout=60)
                    t_sh.append(time.perf_counter() - t1)
            idle[name] = {"image_256_ms": round(float(np.median(t_img[1:])) * 1e3, 3),
                          "shadow_64_angles_ms": round(float(np.median(t_sh[1:])) * 1e3, 3)}
        out["serving_idle"] = dict(idle, what="one request at a time through an idle server (persistent form): a "
                                              "256^2 image of the bench scene and a 64-angle bisection (a=0.9, "
                                              "i=60 deg); in_launch keeps both inside the server's launches")
        print(f"[bench] idle server: {idle}", file=sys.stderr, flush=True)
        del model1, ds1

    if not args.no_heavy:
        from concurrent.futures import ProcessPoolExecutor
        t1 = time.perf_counter()
        with ProcessPoolExecutor(max_workers=8) as ex:
            dss = list(ex.map(_heavy_snapshot, range(8)))
        models8 = [AthenakFluidModel.from_datasets(d, a8, fluid_gamma=13. / 9) for a8, d in dss]  # uploads
        del dss
        print(f"[bench] 8 heavy snapshots ready in {time.perf_counter() - t1:.1f}s", file=sys.stderr, flush=True)
        forms = {"relaunch": [False], "persistent": [True], "both": [False, True]}[args.serve_form]
        for pers in forms:
            hv = heavy_stream(models8, log=sys.stderr, persistent=pers)
            hv["form"] = "persistent launch" if pers else "relaunched every 256 steps"
            out["serving_heavy_persistent" if pers and len(forms) > 1 else "serving_heavy"] = hv
        del models8

    if not args.no_cpu:
        from oracle import c_oracle
        from tests.helpers import c_snapshot_from_model
        cs = c_snapshot_from_model(model)
        nth = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        t1 = time.perf_counter()
        I_cpu, st_cpu = cs.render(a, inc, 1000.0, fov, res, 10000, 13. / 9, 40.0, 1e26, 6.2e9 * MSUN, nu,
                                  nthreads=nth)
        dt_cpu = time.perf_counter() - t1
        out["cpu_baseline"] = {"value": round(n / dt_cpu, 1), "unit": "rays/s", "cores": nth, "kind": "port",
                               "sample": f"the full {res}x{res} image once, C FP64 restatement with OpenMP "
                                         f"({dt_cpu:.1f} s wall)",
                               "flux_rel_err_vs_gpu_plain": abs(float(I_cpu.sum()) - flux_p) / flux_p}
        if "config1_grid" in out:
            # config 1's grid on the host cores: the C FP64 restatement of the
            # same integration; its frozen rows must equal the GPU's
            t1 = time.perf_counter()
            Kc1, _, _ = c_oracle.integrate_final(s1, 10000, 40.0, tol1, a1, nthreads=nth)
            dt1 = time.perf_counter() - t1
            out["cpu_baseline_config1"] = {
                "value": round(n1 / dt1, 1), "unit": "rays/s", "cores": nth, "kind": "port",
                "sample": f"config 1's whole 64x64 grid once, C FP64 restatement with OpenMP ({dt1:.2f} s wall)",
                "frozen_rows_equal_gpu": bool(np.array_equal(Kc1, K1))}
        # the reference's own dataflow: the NumPy FP64 restatement, a scan over
        # steps vectorised over rays (jacfwd by dual numbers, LU inverse), the
        # materialised history and the host meshblock-mask sampler, on one core
        from oracle import mahakala_np as onp
        snap_np = onp.Snapshot(model.all_meshblocks, model.x1v, model.x2v, model.x3v, model.x1f, model.x2f,
                               model.x3f, list(model.variable_names), a, 13. / 9)
        s0 = onp.initialize_geodesics_at_camera(a, inc, 1000, -fov / 2., fov / 2., res)
        idx = np.sort(np.random.Generator(np.random.PCG64(0)).choice(n, 256, replace=False))
        t1 = time.perf_counter()
        S_np, dt_np = onp.geodesic_integrator(10000, s0[idx], 40, 1e-4, a)
        I_np = onp.image_from_history(snap_np, S_np, dt_np, 13. / 9, 1e26, 6.2e9 * MSUN, 40.0, nu)
        dt_np_s = time.perf_counter() - t1
        ref = I_plain.ravel()[idx]
        out["cpu_baseline_numpy"] = {
            "value": round(idx.size / dt_np_s, 2), "unit": "rays/s", "cores": 1, "kind": "port",
            "sample": f"{idx.size} rays drawn at random (seed 0) from the {res}x{res} image, NumPy FP64 "
                      f"restatement with the reference's dataflow, T = {S_np.shape[0]} rows ({dt_np_s:.1f} s)",
            "pixel_max_rel_err_vs_gpu_plain": float(np.max(np.abs(I_np - ref)[ref > 0] / ref[ref > 0]))}
        print(f"[bench] NumPy restatement: {idx.size / dt_np_s:.1f} rays/s", file=sys.stderr, flush=True)
        if "config1_grid" in out:
            # config 1 through the same NumPy dataflow, on 256 of its 4096 rays
            # drawn at random (seed 0); their frozen rows against the GPU's
            idx1 = np.sort(np.random.Generator(np.random.PCG64(0)).choice(n1, 256, replace=False))
            t1 = time.perf_counter()
            _, _, K_np = onp.geodesic_integrator(10000, s1[idx1], 40, tol1, a1, return_steps=True)
            dt_np1 = time.perf_counter() - t1
            out["cpu_baseline_numpy_config1"] = {
                "value": round(idx1.size / dt_np1, 2), "unit": "rays/s", "cores": 1, "kind": "port",
                "sample": f"{idx1.size} of config 1's 4096 rays drawn at random (seed 0), NumPy FP64 restatement "
                          f"with the reference's dataflow ({dt_np1:.1f} s)",
                "frozen_rows_equal_gpu": bool(np.array_equal(K_np, K1[idx1]))}
            print(f"[bench] NumPy restatement, config 1: {idx1.size / dt_np1:.1f} rays/s", file=sys.stderr,
                  flush=True)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
