/*
 * mahakala_hip.h -- C ABI of libmahakala_hip.so, the MI355X (gfx950) kernels
 * behind Mahakala's per-ray hot path.
 *
 * Conventions
 *   - Every entry point returns int: MK_OK (0) or a negative MK_ERR_* code;
 *     mk_last_error() gives the thread's last message.  Nothing throws.
 *   - Pointers named d_* are device pointers (mk_malloc or any HIP
 *     allocation on the current device); h_* are host pointers.  Arrays are
 *     C-contiguous FP64 unless stated.  `stream` is a hipStream_t (NULL for
 *     the default stream); calls are asynchronous on it unless they return
 *     host results, which synchronise it.
 *   - Layouts follow the reference: states (n, 8) = [t, x, y, z, k^t, k^x,
 *     k^y, k^z]; histories (T, n, 8) and (T, n); images pixel p = i*res + j
 *     with image x = grid[i], image y = grid[j] (meshgrid 'ij').
 *   - NaN semantics are the reference's and stay in-band.
 *
 * Each declaration names the reference function it stands in for
 * (file:line in the reference's mahakala/ package).
 */
#ifndef MAHAKALA_HIP_H
#define MAHAKALA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MK_OK 0
#define MK_ERR_ARG (-1)     /* bad argument (shape, range, null pointer)      */
#define MK_ERR_HIP (-2)     /* HIP runtime error                              */
#define MK_ERR_MESH (-3)    /* inconsistent meshblock tree (athenak.py:190)   */
#define MK_ERR_NOMEM (-4)   /* allocation failed or over the memory budget    */
#define MK_ERR_STATE (-5)   /* object in the wrong state (closed server, ...) */
#define MK_ERR_TIMEOUT (-6) /* wait timed out                                 */

/* ------------------------------------------------------------------ runtime */
const char* mk_last_error(void);
int mk_version(void);
/* device name, compute units and HBM size of the current device */
int mk_device_info(char* h_name, size_t name_len, int* h_cus, size_t* h_total_bytes);
/* free and total device memory of the current device (hipMemGetInfo) */
int mk_device_mem(size_t* h_free_bytes, size_t* h_total_bytes);
int mk_malloc(void** d_ptr, size_t bytes);
int mk_free(void* d_ptr);
int mk_memcpy_h2d(void* d_dst, const void* h_src, size_t bytes, void* stream);
int mk_memcpy_d2h(void* h_dst, const void* d_src, size_t bytes, void* stream);
int mk_memset(void* d_dst, int value, size_t bytes, void* stream);
/* device-to-device copy by a shader kernel (16-byte granules); no reference
 * counterpart: used to calibrate HBM traffic counters against known bytes */
int mk_device_copy(void* d_dst, const void* d_src, size_t bytes, void* stream);
/* one wave alone on a SIMD repeats one stage of the per-ray loop `steps`
 * times -> h_out[3] = shader cycles and ns per repetition, and the clock
 * (GHz); no reference counterpart: a measurement aid.  variant: 0/1 RK4
 * step FP64/mixed, 2/3 RK4 + radius + step rule FP64/mixed, 4/5 trilinear
 * interpolation FP64/FP32, 6/7 interpolation + emission FP64/FP32. */
int mk_step_cycles(int variant, int steps, double bhspin, double* h_out);
int mk_stream_create(void** stream);
int mk_stream_destroy(void* stream);
int mk_stream_sync(void* stream);
int mk_event_create(void** event);
int mk_event_destroy(void* event);
int mk_event_record(void* event, void* stream);
int mk_event_sync(void* event);
int mk_event_elapsed_ms(void* start, void* stop, float* h_ms);

/* ------------------------------------------------------------------- camera */
/* get_initial_grid('grid') + initial_condition  (geodesics.py:29-55, 137-181,
 * 204-230): res*res null camera rays -> d_out (res*res, 8). */
int mk_camera_grid(double bhspin, double inclination_deg, double distance, double fov_lower, double fov_upper,
                   int res, double* d_out, void* stream);
/* get_camera_pixel + initial_condition  (geodesics.py:107-134, 219-230) for
 * polar image-plane points; h_radius, h_angle (n,) host. */
int mk_camera_pixels(double bhspin, double inclination_deg, double distance, const double* h_radius,
                     const double* h_angle, long n, double* d_out, void* stream);
/* initial_condition / _Nullify  (geodesics.py:58-85, 219-230): positions
 * d_x (n, 4) and directions d_v (n, 4) -> null states d_out (n, 8). */
int mk_initial_condition(const double* d_x, const double* d_v, long n, double bhspin, double* d_out, void* stream);

/* ------------------------------------------------------------------- metric */
/* metric  (geodesics.py:88-104): d_x (n, 4) -> d_g (n, 4, 4). */
int mk_metric(const double* d_x, long n, double bhspin, double* d_g, void* stream);
/* imetric  (geodesics.py:339-347): closed form eta - f l l -> (n, 4, 4). */
int mk_imetric(const double* d_x, long n, double bhspin, double* d_g, void* stream);
/* rhs / _vectorized_rhs  (geodesics.py:294-314): d_s (n, 8) -> d_out (n, 8). */
int mk_rhs(const double* d_s, long n, double bhspin, double* d_out, void* stream);
/* RK4_gen  (geodesics.py:317-336): d_s (n, 8), d_dt (n,) -> d_out (n, 8). */
int mk_rk4(const double* d_s, const double* d_dt, long n, double bhspin, double* d_out, void* stream);
/* radius_cal  (geodesics.py:284-291): points with `stride` doubles per point
 * (x at offsets 1..3) -> d_r (n,). */
int mk_radius_cal(const double* d_x, long n, int stride, double bhspin, double* d_r, void* stream);

/* --------------------------------------------------------------- integrator */
/* geodesic_integrator  (geodesics.py:233-281), final-state form: per ray the
 * first frozen row K (N if it never froze), the state of its last emitted
 * row (n, 8), and the value select_photons_integrator reads for it
 * (geodesics.py:370-378).  Any output may be NULL. */
int mk_integrate_final(const double* d_s0, long n, int N, double div, double tol, double bhspin, int32_t* d_K,
                       double* d_last, double* d_select_r, void* stream);
/* trim rule (geodesics.py:275-281): T from per-ray K -> *h_T. */
int mk_trim_rows(const int32_t* d_K, long n, int N, long* h_T, void* stream);
/* geodesic_integrator, materialised form: rows 0..T-1 into d_S (T, n, 8)
 * and d_dt (T, n). */
int mk_integrate_history(const double* d_s0, long n, int N, double div, double tol, double bhspin, long T,
                         double* d_S, double* d_dt, void* stream);

/* ----------------------------------------------------------------- shadow */
/* select_photons_integrator  (geodesics.py:354-378): final radius per
 * (radius, angle) camera point; host in, host out. */
int mk_select_photons(double inclination_deg, const double* h_angle, const double* h_radius, long n, double bhspin,
                      double distance, int max_steps, double* h_out);
/* find_shadow_bisection_angles  (geodesics.py:405-435): host in, host out.
 * The radii are the sequential bisection's, bit for bit.  The schedule uses
 * the critical curve of a distant camera as a prediction: up to 4096 angles
 * the first round integrates every midpoint within a band of up to
 * MK_BISECT_BAND (environment, M; default 0.05, 0 = off) about the
 * predicted radius, so most angles finish in one round. */
int mk_shadow_bisect(double bhspin, double inclination_deg, const double* h_angles, long n, int max_steps,
                     double error_allowed, int max_it, double distance, double* h_radii);

/* mk_shadow_bisect without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (request
 * loops that must not call hipMalloc, harnesses that forbid it).  The same
 * arguments and the same radii, bit for bit, as mk_shadow_bisect (in tree
 * rounds: the band schedule's prediction would allocate host memory); the
 * staging of angles and radii lives in a fixed library-owned array of
 * 3 * MK_NOALLOC_MAX_ANGLES doubles (1.5 MiB) per GPU, held until the process
 * ends.  At most MK_NOALLOC_MAX_ANGLES angles (MK_ERR_ARG beyond).  Calls are
 * serialised by a library lock, so they never share the array at once. */
#define MK_NOALLOC_MAX_ANGLES 65536
int mk_shadow_bisect_noalloc(double bhspin, double inclination_deg, const double* h_angles, long n, int max_steps,
                             double error_allowed, int max_it, double distance, double* h_radii);

/* ---------------------------------------------------------------- transfer */
/* synchrotron_coefficients  (transfer.py:30-86), elementwise over n. */
int mk_synchrotron(const double* d_Ne, const double* d_Theta_e, const double* d_B, const double* d_pitch,
                   const double* d_nu, long n, int invariant, double rescale_nu, double* d_j, double* d_alpha,
                   void* stream);
/* solve_specific_intensity  (transfer.py:89-119): (T, npx) inputs ->
 * d_I (npx,); d_dI (T-1, npx) in scan order when non-NULL (dIs=True). */
int mk_solve_intensity(const double* d_j, const double* d_alpha, const double* d_dt, long T, long npx, double L_unit,
                       double* d_I, double* d_dI, void* stream);
/* solve_attenuated_emissivity  (transfer.py:122-144) -> d_out (T-1, npx). */
int mk_attenuated_emissivity(const double* d_j, const double* d_alpha, const double* d_dt, long T, long npx,
                             double L_unit, double* d_out, void* stream);
/* rlow_rhigh_model  (electrons.py:32-50), elementwise. */
int mk_rlow_rhigh(const double* d_dens, const double* d_u, const double* d_beta, long n, double r_low, double r_high,
                  double electron_gamma, double ion_gamma, double* d_Theta_e, void* stream);

/* ---------------------------------------------------------------- snapshot */
typedef struct mk_snapshot mk_snapshot;

/* AthenakFluidModel.load_athenak_meshblocks  (grmhd/athenak.py:67-525) from
 * in-memory datasets: h_prims (8, nmb, n, n, n) float32 holds uov then B in
 * file variable order; h_var_index[8] gives the row of dens, eint, velx,
 * vely, velz, bcc1, bcc2, bcc3 (get_index_for_primitive_by_name,
 * athenak.py:55-65); h_levels (nmb,), h_logical (nmb, 3) as Levels and
 * LogicalLocations; h_x{1,2,3}f (nmb, n+1) and h_x{1,2,3}v (nmb, n).  Builds
 * the ghost-padded FP32 layout on the device (same-level copy, coarser
 * injection, finer 2x2x2 restriction) and an O(1) leaf lookup grid. */
int mk_snapshot_create(const float* h_prims, const int32_t* h_var_index, long nmb, int n, const int32_t* h_levels,
                       const int32_t* h_logical, const double* h_x1f, const double* h_x2f, const double* h_x3f,
                       const double* h_x1v, const double* h_x2v, const double* h_x3v, mk_snapshot** out);
int mk_snapshot_free(mk_snapshot* snap);
/* device bytes held, and the ghost-padded cells as (nmb, n+2, n+2, n+2, 8)
 * float32 in canonical order [rho, u, U1, U2, U3, B1, B2, B3] */
int mk_snapshot_bytes(const mk_snapshot* snap, size_t* h_bytes);
int mk_snapshot_download(const mk_snapshot* snap, float* h_cells);
/* meshblock id per point (athenak.py:657-670), -1 outside; d_S (n, stride) */
int mk_block_index(const mk_snapshot* snap, const double* d_S, long n, int stride, int32_t* d_mb, void* stream);
/* get_prims_from_geodesics  (athenak.py:527-637): d_S (n, 8) -> d_out (8, n)
 * in canonical order, zero outside the domain. */
int mk_sample_prims(const mk_snapshot* snap, const double* d_S, long n, double* d_out, void* stream);
/* get_fluid_scalars_from_geodesics  (athenak.py:639-812): d_S (n, 8) ->
 * d_out (5, n) = dens, u, pitch_angle, kdotu, b. */
int mk_sample_scalars(const mk_snapshot* snap, const double* d_S, long n, double bhspin, double fallback_pitch_angle,
                      double* d_out, void* stream);

/* ----------------------------------------------------------- fused render */
typedef struct mk_render_params {
  double bhspin;              /* a (overrides the snapshot's spin)          */
  double inclination;         /* camera inclination, degrees                */
  double distance;            /* camera distance, GM/c^2                    */
  double fov;                 /* full field of view, GM/c^2                 */
  double mass_scale;          /* g                                          */
  double M_bh;                /* g                                          */
  double r_high;              /* R(beta) model                              */
  double fluid_gamma;         /* adiabatic index of the fluid               */
  double observing_frequency; /* Hz                                         */
  double div;                 /* step divisor (40)                          */
  double tol;                 /* horizon tolerance (1e-4)                   */
  double fallback_pitch;      /* pi/3                                       */
  double sigma_cut;           /* 100                                        */
  double fp64_radius;         /* mixed: FP64 steps where r < this           */
  double ring_radius;         /* mixed: FP64 where r < this and             */
  double ring_drdl;           /*        |dr/dlambda| < this; an 8x8 tile's  */
                              /* rays step in FP64 while any ray is there   */
  int32_t resolution;         /* pixels per side                            */
  int32_t max_nsteps;         /* N (10000)                                  */
  int32_t precision;          /* 0 = plain FP64, 1 = mixed FP32/FP64        */
  int32_t flags;              /* MK_RENDER_* bits, 0 by default             */
} mk_render_params;

/* mk_render_params.flags: kernel form.  By default the library runs a launch
 * of up to 3 8x8 tiles per SIMD in the triple form -- a geodesic wave, a
 * gather wave (block lookup, cell loads, interpolation) and an emission wave
 * per tile, with the snapshot's block tables in LDS where they fit and in
 * global memory otherwise -- and a larger one in the classic form (one wave
 * per tile).  The paired form (a geodesic wave and a transfer wave per tile)
 * runs only when forced and its tables fit in LDS (else the classic form).
 * The bits force one form; all give the same image, bit for bit. */
#define MK_RENDER_CLASSIC 1
#define MK_RENDER_PAIRED 2
#define MK_RENDER_TRIPLE 4
/* Testing aid for the paired and triple forms: every bounded wait between
 * the waves of a tile gives up after one poll, so the failure path runs (a
 * wave that gives up marks its ring failed, and the tile's pixels are NaN
 * and its step counts -1).  Normal waits give up after 2^24 polls. */
#define MK_RENDER_WAIT_TEST 256
/* Hand tiles out in row-major order instead of photon ring first (the same
 * image, bit for bit; only the schedule differs). */
#define MK_RENDER_ROW_MAJOR 512
/* Classic form: by default the launch holds only the workgroups resident at
 * once and each wave takes the next tile from a counter until none is left;
 * this bit launches one wave per tile instead (the same image, bit for bit). */
#define MK_RENDER_STATIC_TILES 1024

/* make_image  (images.py:30-144) for a device-resident snapshot: camera,
 * geodesic, sampling, emission and transfer fused per ray; d_image
 * (res*res,) FP64, d_steps (res*res,) accepted steps per ray (may be NULL). */
int mk_render(const mk_snapshot* snap, const mk_render_params* params, double* d_image, int32_t* d_steps,
              void* stream);
/* the kernel form mk_render runs for these parameters -> *h_form
 * (MK_RENDER_CLASSIC, MK_RENDER_PAIRED or MK_RENDER_TRIPLE); no reference
 * counterpart */
int mk_render_form(const mk_snapshot* snap, const mk_render_params* params, int32_t* h_form);

/* ---------------------------------------------------------------- serving */
/* Continuous batching of image and shadow requests on one GPU.  No single
 * reference counterpart: the reference renders one image per make_image
 * call (images.py:30-144) and one bisection per find_shadow_bisection_angles
 * call (geodesics.py:405-435); the server runs both kinds of work, for many
 * requests of different size, spin, inclination and frequency, in shared
 * launches of packed rays.  A worker thread admits queued requests while
 * their output buffers fit the HBM budget, and launches an advance kernel in
 * which every wave runs up to steps_per_launch RK4 steps on a tile of 64
 * rays or bisection angles of one request and, when the tile is done, takes
 * the next tile of any admitted request.  Results do not depend on what else
 * is in flight. */
typedef struct mk_server mk_server;

typedef struct mk_server_config {
  int32_t precision;        /* image rays: 0 plain FP64, 1 mixed (bisection is FP64) */
  int32_t lanes;            /* ray slots = lanes of the advance kernel; 0 = 2 waves per SIMD */
  int32_t steps_per_launch; /* RK4 steps per lane per launch; 0 = 256 */
  int32_t max_requests;     /* requests in flight; 0 = 1024 */
  size_t hbm_budget;        /* device bytes for in-flight outputs; 0 = 4 GiB */
  int32_t use_graph;        /* replay the per-iteration launch as a hipGraph */
  int32_t flags;            /* MK_SERVER_* bits, 0 by default */
} mk_server_config;

/* mk_server_config.flags: one persistent launch instead of a launch per
 * steps_per_launch steps.  Its waves take tiles in the order the worker
 * publishes them, run each tile to the end and stay resident while work keeps
 * coming; the launch ends once the server has been idle for 20 ms (or, after
 * 1 s, whenever nothing is in flight) and the next request relaunches it.
 * A launch with no bisection to serve runs a kernel without the bisection
 * path, at three waves per SIMD (mixed, default lanes) when it starts with a
 * deep backlog; a launch starting from idle first gathers work for 0.2 ms,
 * and a bisection arriving during an image-only launch waits until that
 * launch has drained.  Work, descriptors and outputs pass through pinned
 * host memory, and hbm_budget (default 1 GiB here) bounds the pinned output
 * staging.  The results are the same bits as the relaunched form's and
 * mk_render's.  Statistics of this form are updated when a launch ends. */
#define MK_SERVER_PERSISTENT 1
/* mk_server_config.flags: bisections always run inside the server's
 * launches.  Without it, an idle server (no launch running, nothing in
 * flight) runs a bisection of up to 1024 angles at the head of its queue
 * itself, in mk_shadow_bisect's band schedule: about one integration chain
 * instead of one per halving (64 angles: 3.5 against 35 ms); the same radii,
 * bit for bit.  Requests submitted meanwhile wait until it ends. */
#define MK_SERVER_SHADOW_IN_LAUNCH 2
/* mk_server_config.flags: images always run inside the server's launches.
 * Without it, an idle server whose queue holds one image of up to 512^2 and
 * nothing else renders it itself through mk_render (up to 256^2 three waves
 * per tile).  The same image, bit for bit.  Above 256^2 that call builds
 * mk_render's tile order for the view, as mk_render's own first call does. */
#define MK_SERVER_IMAGE_IN_LAUNCH 4

typedef struct mk_image_request {
  int32_t snapshot;         /* id from mk_server_add_snapshot */
  int32_t resolution;
  int32_t max_nsteps;       /* 10000 in make_image */
  int32_t priority;         /* higher is admitted first; equal: submit order (0 by default) */
  double bhspin, inclination, distance, fov, observing_frequency, mass_scale, M_bh, r_high, fluid_gamma;
} mk_image_request;

typedef struct mk_shadow_request {
  double bhspin, inclination, distance, error_allowed;
  int32_t n_angles, max_steps, max_it, priority; /* priority as for images */
  const double* h_angles;   /* copied at submit */
} mk_shadow_request;

typedef struct mk_server_stats {
  int64_t submitted, admitted, completed, queued, in_flight;
  int64_t items_done;       /* image rays + bisection angles finished */
  int64_t iterations;       /* advance launches */
  double gpu_ms;            /* summed launch time (HIP events) */
  size_t hbm_in_use, hbm_budget;
  int64_t lane_steps_busy;  /* lane-steps that advanced a ray or bisection */
  int64_t lane_steps_total; /* lanes x steps_per_launch x launches; persistent
                               form: 64 x the wave-steps of every tile run */
  int32_t lanes, steps_per_launch; /* effective configuration */
} mk_server_stats;

int mk_server_create(const mk_server_config* cfg, mk_server** out);
/* register a resident snapshot (not copied; keep it alive) -> *h_id */
int mk_server_add_snapshot(mk_server* srv, const mk_snapshot* snap, int32_t* h_id);
/* queue an image; h_out (resolution^2 doubles, pixel i*res+j) is written on
 * completion; MK_ERR_NOMEM if its footprint can never fit the budget */
int mk_server_submit_image(mk_server* srv, const mk_image_request* req, double* h_out, int64_t* h_id);
/* queue a shadow bisection; h_out (n_angles doubles) written on completion */
int mk_server_submit_shadow(mk_server* srv, const mk_shadow_request* req, double* h_out, int64_t* h_id);
/* completed requests since the last poll: ids and submit-to-done latency
 * (the newest 2^20 completions are kept between polls) */
int mk_server_poll(mk_server* srv, int64_t* h_ids, double* h_latency_ms, int32_t cap, int32_t* h_n);
/* block until request id (or all, id < 0) completes; MK_ERR_TIMEOUT after
 * timeout_s seconds, MK_ERR_ARG for an id never issued */
int mk_server_wait(mk_server* srv, int64_t id, double timeout_s);
/* Removes a request that is still queued (not yet admitted to a launch nor
 * being run): *h_cancelled = 1, its output is never written, and a wait on it
 * returns MK_ERR_STATE; otherwise *h_cancelled = 0 (it is running or done).
 * No counterpart in the reference. */
int mk_server_cancel(mk_server* srv, int64_t id, int32_t* h_cancelled);
int mk_server_get_stats(mk_server* srv, mk_server_stats* h_stats);
int mk_server_destroy(mk_server* srv);

#ifdef __cplusplus
}
#endif
#endif /* MAHAKALA_HIP_H */
