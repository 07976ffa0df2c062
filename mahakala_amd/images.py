"""Images of a fluid model (reference: mahakala/images.py:30-144).

``make_image`` keeps the reference's signature and adds ``precision``
('plain64' by default, or 'mixed').  For a model with a resident snapshot
(``AthenakFluidModel``) it runs the fused per-ray kernel: camera ray,
geodesic, sampling, emission and transfer in one launch, without the
(T, npx, 8) history.  For any other fluid model it follows the reference's
chunked pipeline with the GPU integrator, the model's own
``get_fluid_scalars_from_geodesics`` and the GPU transfer kernels.
"""
from __future__ import annotations

import numpy as np

from mahakala_amd import _lib
from mahakala_amd.constants import Msun
from mahakala_amd.electrons import rlow_rhigh_model

FORMS = {None: 0, "classic": 1, "paired": 2, "triple": 4}  # MK_RENDER_CLASSIC, _PAIRED, _TRIPLE
PRECISIONS = {"plain64": 0, "fp64": 0, "mixed": 1, "mixed-zone": 1}

# Mixed precision keeps in FP64 the ray state and its RK4 accumulation, the
# step-size / termination rule, the camera-ray null normalisation and the
# optical-depth / intensity accumulation; the RK4 right-hand sides, sampling
# and emission run in FP32.  "mixed-zone" also takes the RK4 steps in FP64
# inside r < 4 M and near the photon ring (r < 5 M with |dr/dlambda| < 0.1):
# an 8x8 tile's rays step in FP64 while any of them is in that zone.
# On the 256^2 torus image the zone changes the total flux by < 1e-7
# relative and costs ~30% in time, so plain "mixed" leaves it off.
ZONES = {"mixed": (0.0, 0.0, 0.0), "mixed-zone": (4.0, 5.0, 0.1)}

# Device bytes per (row, ray) of the unfused pipeline at its peak, the
# sampler: the states in (8 FP64) and five scalars out, rounded up.
UNFUSED_ROW_BYTES = 128


def render_params(bhspin, camera_inclination=60, camera_distance=1000, mass_scale=1.e26, M_bh=6.2e9 * Msun,
                  r_high=40, observing_frequency=230.e9, fov=20, resolution=160, max_nsteps=10000,
                  fluid_gamma=13. / 9, precision="plain64", fp64_radius=None, ring_radius=None,
                  ring_drdl=None, form=None):
    """mk_render parameters.  form: None (the library chooses), "classic"
    (one wave per 8x8 tile), "paired" (a geodesic wave and a transfer wave
    per tile) or "triple" (geodesic, gather and emission waves per tile);
    all forms give the same image, bit for bit."""
    if precision not in PRECISIONS:
        raise ValueError(f"precision must be one of {sorted(PRECISIONS)}")
    if form not in FORMS:
        raise ValueError(f"form must be one of {sorted(FORMS, key=str)}")
    p = _lib.mk_render_params()
    p.bhspin = float(bhspin)
    p.inclination = float(camera_inclination)
    p.distance = float(camera_distance)
    p.fov = float(fov)
    p.mass_scale = float(mass_scale)
    p.M_bh = float(M_bh)
    p.r_high = float(r_high)
    p.fluid_gamma = float(fluid_gamma)
    p.observing_frequency = float(observing_frequency)
    p.div = 40.0
    p.tol = 1e-4
    p.fallback_pitch = np.pi / 3.
    p.sigma_cut = 100.0
    zr, zring, zdrdl = ZONES.get(precision, (0.0, 0.0, 0.0))
    p.fp64_radius = zr if fp64_radius is None else float(fp64_radius)
    p.ring_radius = zring if ring_radius is None else float(ring_radius)
    p.ring_drdl = zdrdl if ring_drdl is None else float(ring_drdl)
    p.resolution = int(resolution)
    p.max_nsteps = int(max_nsteps)
    p.precision = PRECISIONS[precision]
    p.flags = FORMS[form]
    return p


def render(snapshot, params, return_steps=False):
    """Fused image kernel on a resident snapshot -> (res, res) FP64 [, steps]."""
    res = params.resolution
    n = res * res
    img = _lib.DeviceBuffer(8 * n)
    steps = _lib.DeviceBuffer(4 * n) if return_steps else None
    _lib.check(_lib.lib().mk_render(snapshot.handle, params, img.ptr, steps.ptr if steps else None, None))
    I = img.to_array((res, res))
    if return_steps:
        return I, steps.to_array((res, res), np.int32)
    return I


def make_image(fluid_model, camera_inclination=60, camera_distance=1000, mass_scale=1.e26, M_bh=6.2e9 * Msun,
               r_high=40, observing_frequency=230.e9, fov=20, resolution=160, max_nsteps=10000,
               max_chunk_bytes=None, precision="plain64"):
    """Image of specific intensities (res, res) in cgs  (images.py:30-144)."""
    import mahakala_amd as ma

    bhspin = fluid_model.bhspin
    fluid_gamma = fluid_model.fluid_gamma
    npx = int(resolution) ** 2
    snap = getattr(fluid_model, "device_snapshot", None)
    num_pixels_per_chunk = npx + 10
    if max_chunk_bytes is not None:
        num_pixels_per_chunk = max(1, int(max_chunk_bytes // 4 // 20 // max_nsteps))
    elif snap is None:
        # no budget given: chunks whose per-step arrays fit half of the free
        # device memory at the longest possible history (max_nsteps rows)
        free, _ = _lib.device_memory()
        num_pixels_per_chunk = max(1, min(npx + 10, int(free // 2 // UNFUSED_ROW_BYTES // max_nsteps)))
    print(num_pixels_per_chunk, npx)  # the reference prints the same pair (images.py:76)
    if fluid_gamma is None:
        raise TypeError("fluid_model.fluid_gamma must be set to compute beta (images.py:88)")

    if snap is not None:
        p = render_params(bhspin, camera_inclination, camera_distance, mass_scale, M_bh, r_high,
                          observing_frequency, fov, resolution, max_nsteps, fluid_gamma, precision)
        return render(snap, p)

    if precision != "plain64":
        raise ValueError("mixed precision needs a resident snapshot (AthenakFluidModel)")
    s0 = ma.initialize_geodesics_at_camera(bhspin, camera_inclination, camera_distance, -fov / 2., fov / 2.,
                                           resolution)
    units = fluid_model.get_units(M_bh, mass_scale)
    out = []
    lo = 0
    while lo < s0.shape[0]:
        chunk = s0[lo:lo + num_pixels_per_chunk]
        S, final_dt = ma.geodesic_integrator(max_nsteps, chunk, 40, 1e-4, bhspin)
        fs = fluid_model.get_fluid_scalars_from_geodesics(S)
        with np.errstate(divide="ignore", invalid="ignore"):
            bsq = fs['b'] * fs['b']
            beta = fs['u'] * (fluid_gamma - 1.) / bsq / 0.5
            sigma = bsq / fs['dens']
        Theta_e = rlow_rhigh_model(fs['dens'], fs['u'], beta, r_high=r_high)
        Ne = units['Ne_unit'] * fs['dens']
        Bg = units['B_unit'] * fs['b']
        local_nu = - fs['kdotu'] * observing_frequency
        em, ab = ma.synchrotron_coefficients(Ne, Theta_e, Bg, fs['pitch_angle'], local_nu, invariant=True,
                                             rescale_nu=1. / observing_frequency)
        with np.errstate(invalid="ignore"):
            cut = np.asarray(sigma) > 100.
        em = em.at[cut].set(0.)
        ab = ab.at[cut].set(0.)
        out.append(np.asarray(ma.solve_specific_intensity(em, ab, final_dt, units['L_unit'])))
        lo += num_pixels_per_chunk
    return np.concatenate(out).reshape((resolution, resolution))
