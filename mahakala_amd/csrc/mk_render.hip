// mk_render.hip -- fused per-ray image kernel: camera ray, RK4 null geodesic
// in Cartesian Kerr-Schild, meshblock sampling, thermal synchrotron
// emission/absorption and transfer, one ray per lane (images.py:30-144 of
// the reference, with geodesics.py, athenak.py, electrons.py, transfer.py).
//
// The reference materialises the (T, npx, 8) history, samples every row and
// scans the explicit-Euler transfer backward from the far end:
//     I <- I - dt[i-1] L (j[i] - alpha[i] I),  i = T-1 .. 1.
// That recurrence is linear in I, so with c_i = -dt[i-1] L it equals
//     I = sum_{i>=1} c_i j_i prod_{k<i} (1 - c_k alpha_k),
// which this kernel accumulates in the order the ray is integrated, keeping
// no history.  Rows i > K (frozen) have dt = 0 and add nothing; row 0 never
// contributes; row N is never emitted, so the sample after step N-1 is
// skipped.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>
#include <vector>

#include "mk_device.h"
#include "mk_internal.h"
#include "mk_render.h"
#include "mk_snapshot.h"

namespace mk {

// One ray: the sample of accepted step k is located (lookup tables in LDS)
// and its 8 corner records are requested right after the step, then
// consumed after the next RK4 step has been computed, so the gather latency
// overlaps the geodesic arithmetic.  Contributions are still added in step
// order, so the result is the same as sampling in place.
// CACHE (tables in global memory): the ray keeps its last block record and
// skips the lookup while its samples stay in that block.  IN_PLACE (mixed
// precision, tables in LDS, three workgroups per CU): no pipeline; the sample is consumed
// right after its step, so no corner records are held across the next RK4
// step and the kernel fits three waves per SIMD (168 VGPRs), which hides the
// load latency better than the pipeline does at two.
template <bool MIXED, bool CACHE, bool IN_PLACE>
__device__ __forceinline__ void trace_ray(const SnapView& sv, const double* meta, const int* grid, const RenderDev& rp,
                                          double x_img, double y_img, double& I_out, int& steps_out) {
  using T = typename std::conditional<MIXED, float, double>::type;
  double sx[4], sv4[4], s[8];
  camera_ray(x_img, y_img, rp.ci, rp.si, rp.dist, sx, sv4);
  nullify(sx, sv4, rp.a, s);
  double r = radius_cal(s[1], s[2], s[3], rp.a);
  double dt = step_rule(r, rp.rEH, rp.div, rp.tol);
  double I = 0.0, trans = 1.0, drdl = 1e300;
  int nacc = 0;
  bool pend = false;
  Stencil pst;
  Corners pc;
  double ps[8], pdt = 0.0;
  int cmb = -1;
  double crec[kMeta];
  // Outside the domain (all primitives zero) and wherever the density is
  // zero, the reference's j and alpha come out 0 (or NaN, then 0), and the
  // step adds exactly nothing: those samples skip the emission arithmetic.
  auto consume = [&]() {
    if (pst.mb < 0) return;
    T P[8], j, al;
    interp<T>(pst, pc, P);
    if (P[0] == T(0)) return;
    emission_from_prims<T>(rp.ip, P, ps, rp.a, j, al);
    const double c = -pdt * rp.ip.L_unit;
    I += trans * (c * (double)j);
    trans *= (1.0 - c * (double)al);
  };
  // The same sum without branches: lanes with nothing to add evaluate the
  // emission too and keep I and the transmission by selection.
  auto consume_flat = [&]() {
    T P[8], j, al;
    interp_raw<T>(pst, pc, P);
    emission_from_prims<T>(rp.ip, P, ps, rp.a, j, al);
    const bool use = pend && pst.mb >= 0 && P[0] != T(0);
    const double c = -pdt * rp.ip.L_unit;
    const double In = I + trans * (c * (double)j);
    const double tn = trans * (1.0 - c * (double)al);
    I = use ? In : I;
    trans = use ? tn : trans;
  };
  const bool zone = MIXED && (rp.fp64_r > 0.0 || rp.ring_r > 0.0);
#pragma unroll
  for (int q = 0; q < 8; ++q) pc.q[q][0] = pc.q[q][1] = make_float4(0.f, 0.f, 0.f, 0.f);
  pst.mb = -1;
  pst.base = 0;
  pst.fr[0] = pst.fr[1] = pst.fr[2] = 0.0;
  for (int k = 0; k < rp.N; ++k) {
    if (dt == 0.0) break;
    double nw[8];
    if (!zone && __any(pend && pst.mb >= 0)) {
      // a sample of the wave is inside the domain: the next RK4 step and
      // that sample's emission are independent, and one straight block lets
      // a wave that is alone on its SIMD interleave the two
      if constexpr (!MIXED) rk4_step<double>(s, dt, rp.a, nw);
      else rk4_step_mixed(s, dt, (float)rp.a, nw);
      consume_flat();
    } else {
      if constexpr (!MIXED) {
        rk4_step<double>(s, dt, rp.a, nw);
      } else {
        const bool fp64 = zone && __any((r < rp.fp64_r) || (r < rp.ring_r && __builtin_fabs(drdl) < rp.ring_drdl));
        if (fp64) rk4_step<double>(s, dt, rp.a, nw);
        else rk4_step_mixed(s, dt, (float)rp.a, nw);
      }
      if (pend) consume();
    }
    pend = false;
    double rn, dtn;
    if constexpr (!MIXED) {
      rn = radius_cal(nw[1], nw[2], nw[3], rp.a);
      dtn = step_rule(rn, rp.rEH, rp.div, rp.tol);
    } else {
      rn = radius_f32(nw[1], nw[2], nw[3], rp.a);
      dtn = step_rule_fast(rn, rp.rEH, rp.inv_div, rp.div, rp.tol);
    }
    if (dtn == 0.0) break;
    ++nacc;
    if (k + 1 < rp.N) {  // row N is never emitted
      if constexpr (CACHE) pst = locate_cached(sv, meta, grid, nw[1], nw[2], nw[3], cmb, crec);
      else pst = locate(sv, meta, grid, nw[1], nw[2], nw[3]);
      fetch(sv, pst, pc);
#pragma unroll
      for (int q = 0; q < 8; ++q) ps[q] = nw[q];
      pdt = dt;
      pend = true;
      if constexpr (IN_PLACE) {
        consume();
        pend = false;
      }
    }
    if (zone) drdl = (rn - r) / dt;
#pragma unroll
    for (int q = 0; q < 8; ++q) s[q] = nw[q];
    dt = dtn;
    r = rn;
  }
  if (pend) consume();
  I_out = I;
  steps_out = nacc;
}

// Pixel of lane: each wave takes an 8x8 tile; workgroup ids are remapped so
// each of the 8 XCDs (hardware assigns workgroups round-robin) walks a
// contiguous band of tiles and shares snapshot lines in its own L2.
__device__ __forceinline__ bool pixel_of_lane(const RenderDev& rp, int& i, int& j) {
  const unsigned nwg = gridDim.x, wg = blockIdx.x;
  const unsigned per = nwg / 8;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  long tile;
  if (rp.order) {  // workgroups start in index order: the longest tiles first
    const long k = (long)wg * (blockDim.x >> 6) + wave;
    if (k >= rp.ntiles) return false;
    tile = rp.order[k];
  } else {
    const unsigned lwg = (nwg % 8 == 0) ? (wg % 8) * per + wg / 8 : wg;
    tile = (long)lwg * (blockDim.x >> 6) + wave;
  }
  if (tile >= rp.ntiles) return false;
  const int ti = (int)(tile / rp.tiles_x), tj = (int)(tile % rp.tiles_x);
  i = ti * 8 + (lane >> 3);
  j = tj * 8 + (lane & 7);
  return i < rp.res && j < rp.res;
}

// Copies the block geometry and the leaf lookup grid into LDS (when they
// fit), so block lookup and stencil arithmetic never wait on HBM.
template <bool LDS>
__device__ __forceinline__ void stage_tables(const SnapView& sv, double* lds, const double*& meta, const int*& grid) {
  meta = sv.meta;
  grid = sv.grid;
  if constexpr (LDS) {
    const int nm = sv.nmb * kMeta, ng = sv.gx * sv.gy * sv.gz;
    for (int t = threadIdx.x; t < nm; t += blockDim.x) lds[t] = sv.meta[t];
    int* g = reinterpret_cast<int*>(lds + nm);
    for (int t = threadIdx.x; t < ng; t += blockDim.x) g[t] = sv.grid[t];
    __syncthreads();
    meta = lds;
    grid = g;
  }
}

// IN_PLACE (mixed, tables in LDS small enough for three workgroups per CU):
// samples consumed right after their step, three waves per SIMD.
//
// With rp.next_tile the grid holds only the workgroups resident at once: each
// stages the tables once, and each of its waves takes the next tile (in tile
// order) from the counter until none is left, so a wave whose tile ends early
// starts another instead of idling until its workgroup's longest tile ends.
// next_tile[1] counts the waves that found none left; the last of them sets
// both words back to 0, so the next launch on the stream starts from 0.
template <bool MIXED, bool LDS, bool IN_PLACE>
__global__ void __launch_bounds__(256, IN_PLACE ? 3 : 2) k_render(SnapView sv, RenderDev rp, double* __restrict__ img,
                                                                  int32_t* __restrict__ steps) {
  extern __shared__ double lds[];
  const double* meta;
  const int* grid;
  stage_tables<LDS>(sv, lds, meta, grid);
  if (rp.next_tile) {
    const int lane = threadIdx.x & 63;
    while (true) {
      int k = 0;
      if (lane == 0) k = atomicAdd(rp.next_tile, 1);
      k = __builtin_amdgcn_readfirstlane(__shfl(k, 0));
      if (k >= rp.ntiles) {
        // every other wave has taken its last ticket before it counts itself
        // here, so the last one to count leaves no taker behind
        if (lane == 0 && atomicAdd(rp.next_tile + 1, 1) == (int)(gridDim.x * (blockDim.x >> 6)) - 1) {
          __hip_atomic_store(rp.next_tile, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          __hip_atomic_store(rp.next_tile + 1, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        break;
      }
      const long tile = rp.order ? rp.order[k] : k;
      const int i = (int)(tile / rp.tiles_x) * 8 + (lane >> 3), j = (int)(tile % rp.tiles_x) * 8 + (lane & 7);
      if (i < rp.res && j < rp.res) {
        double I;
        int n;
        trace_ray<MIXED, !LDS, IN_PLACE>(sv, meta, grid, rp, grid_coord(rp.lo, rp.step, i),
                                         grid_coord(rp.lo, rp.step, j), I, n);
        const long p = (long)i * rp.res + j;
        img[p] = I;
        if (steps) steps[p] = n;
      }
    }
    return;
  }
  int i, j;
  if (!pixel_of_lane(rp, i, j)) return;
  double I;
  int n;
  trace_ray<MIXED, !LDS, IN_PLACE>(sv, meta, grid, rp, grid_coord(rp.lo, rp.step, i), grid_coord(rp.lo, rp.step, j),
                                   I, n);
  const long p = (long)i * rp.res + j;
  img[p] = I;
  if (steps) steps[p] = n;
}

// --------------------------------------------------------------- paired form
// For launches of at most one tile per SIMD (a 256^2 image is 1024 tiles on
// 1024 SIMDs) the time is the slowest tile's step count times the step
// latency of a wave alone on its SIMD.  The paired form gives each tile two
// waves: a geodesic wave (camera ray, RK4, step rule) and a transfer wave
// (block lookup, gather, interpolation, emission, transfer).  They run
// concurrently and hand samples over through a ring of kPairDepth slots in
// LDS, synchronised by two counters per pair (rounds produced, rounds
// consumed) with release/acquire ordering at workgroup scope; no barrier
// couples different tiles.  Waits are bounded: after kSpinLimit polls a wave
// gives up and writes NaN for its pixels, so every wave exits.  The transfer
// wave adds the samples in step order with the same expressions as the
// classic form.
constexpr int kPairTiles = 4;   // tiles per workgroup (geodesic + transfer wave each)
constexpr int kPairDepth = 4;   // ring slots per tile
constexpr int kPairComp = 9;    // state[8], dt
constexpr size_t kPairRingDoubles = (size_t)kPairTiles * kPairDepth * kPairComp * 64;
constexpr size_t kPairLdsExtra = kPairRingDoubles * sizeof(double) + kPairTiles * 4 * sizeof(int);
constexpr int kSpinLimit = 1 << 24;

// Polls *cnt (LDS) until it reaches target; false after `limit` polls.
__device__ __forceinline__ bool wait_at_least(int* cnt, int target, int limit) {
  for (int spin = 0; spin < limit; ++spin) {
    const int v = __builtin_amdgcn_readfirstlane(
        __hip_atomic_load(cnt, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP));
    if (v >= target) return true;
    __builtin_amdgcn_s_sleep(1);
  }
  return false;
}

__device__ __forceinline__ void publish(int* cnt, int v) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  if ((threadIdx.x & 63) == 0) __hip_atomic_store(cnt, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

template <bool MIXED>
__global__ void __launch_bounds__(128 * kPairTiles) k_render_pair(SnapView sv, RenderDev rp, double* __restrict__ img,
                                                                  int32_t* __restrict__ steps, int tab_doubles) {
  using T = typename std::conditional<MIXED, float, double>::type;
  extern __shared__ double lds[];
  double* ring_all = lds + tab_doubles;
  int* ctr = reinterpret_cast<int*>(ring_all + kPairRingDoubles);  // per tile: produced, consumed, done, pad
  if (threadIdx.x < 4 * kPairTiles) ctr[threadIdx.x] = ((threadIdx.x & 3) == 2) ? -1 : 0;
  const double* meta;
  const int* grid;
  stage_tables<true>(sv, lds, meta, grid);  // ends with a workgroup barrier
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int pair = wave % kPairTiles;
  const bool geo = wave < kPairTiles;
  double* ring = ring_all + (size_t)pair * kPairDepth * kPairComp * 64 + lane;
  int* produced = ctr + 4 * pair;
  int* consumed = produced + 1;
  int* done = produced + 2;
  const unsigned nwg = gridDim.x, wg = blockIdx.x, per = nwg / 8;
  const unsigned lwg = (nwg % 8 == 0) ? (wg % 8) * per + wg / 8 : wg;
  const long tile = (long)lwg * kPairTiles + pair;
  int pi = 0, pj = 0;
  bool valid = tile < rp.ntiles;
  if (valid) {
    pi = (int)(tile / rp.tiles_x) * 8 + (lane >> 3);
    pj = (int)(tile % rp.tiles_x) * 8 + (lane & 7);
    valid = pi < rp.res && pj < rp.res;
  }
  const long p = (long)pi * rp.res + pj;
  if (geo) {
    double s[8], r = 0.0, dt = 0.0, drdl = 1e300;
    int k = 0, nacc = 0;
    bool alive = false;
    if (valid) {
      double sx[4], sv4[4];
      camera_ray(grid_coord(rp.lo, rp.step, pi), grid_coord(rp.lo, rp.step, pj), rp.ci, rp.si, rp.dist, sx, sv4);
      nullify(sx, sv4, rp.a, s);
      r = radius_cal(s[1], s[2], s[3], rp.a);
      dt = step_rule(r, rp.rEH, rp.div, rp.tol);
      alive = true;
    }
    const bool zone = MIXED && (rp.fp64_r > 0.0 || rp.ring_r > 0.0);
    bool ok = true;
    int t = 0;
    while (__any(alive)) {
      if (!wait_at_least(consumed, t - kPairDepth + 1, rp.spin_limit)) {
        ok = false;
        break;
      }
      double* out = ring + (size_t)(t % kPairDepth) * kPairComp * 64;
      bool wrote = false;
      if (alive) {
        if (k >= rp.N || dt == 0.0) {
          alive = false;
        } else {
          double nw[8];
          if constexpr (!MIXED) {
            rk4_step<double>(s, dt, rp.a, nw);
          } else {
            const bool fp64 = zone && __any((r < rp.fp64_r) || (r < rp.ring_r && __builtin_fabs(drdl) < rp.ring_drdl));
            if (fp64) rk4_step<double>(s, dt, rp.a, nw);
            else rk4_step_mixed(s, dt, (float)rp.a, nw);
          }
          double rn, dtn;
          if constexpr (!MIXED) {
            rn = radius_cal(nw[1], nw[2], nw[3], rp.a);
            dtn = step_rule(rn, rp.rEH, rp.div, rp.tol);
          } else {
            rn = radius_f32(nw[1], nw[2], nw[3], rp.a);
            dtn = step_rule_fast(rn, rp.rEH, rp.inv_div, rp.div, rp.tol);
          }
          if (dtn == 0.0) {
            alive = false;
          } else {
            ++nacc;
            if (k + 1 < rp.N) {  // row N is never emitted
#pragma unroll
              for (int c = 0; c < 8; ++c) out[c * 64] = nw[c];
              out[8 * 64] = dt;
              wrote = true;
            }
            if (zone) drdl = (rn - r) / dt;
#pragma unroll
            for (int c = 0; c < 8; ++c) s[c] = nw[c];
            dt = dtn;
            r = rn;
            if (++k >= rp.N) alive = false;
          }
        }
      }
      if (!wrote) out[8 * 64] = 0.0;
      publish(produced, ++t);
    }
    publish(done, ok ? t : -2);
    if (valid && steps) steps[p] = ok ? nacc : -1;
    return;
  }
  // transfer wave
  double I = 0.0, trans = 1.0, ps[8], pdt = 0.0;
  bool pend = false, ok = true;
  Stencil pst;
  Corners pc;
  auto consume = [&]() {
    if (pst.mb < 0) return;
    T P[8], j, al;
    interp<T>(pst, pc, P);
    if (P[0] == T(0)) return;
    emission_from_prims<T>(rp.ip, P, ps, rp.a, j, al);
    const double c = -pdt * rp.ip.L_unit;
    I += trans * (c * (double)j);
    trans *= (1.0 - c * (double)al);
  };
  for (int t = 0;; ++t) {
    // round t is ready, or the geodesic wave has finished after fewer rounds
    int state = 0;
    for (int spin = 0; spin < rp.spin_limit; ++spin) {
      const int pr = __builtin_amdgcn_readfirstlane(
          __hip_atomic_load(produced, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP));
      if (pr > t) {
        state = 1;
        break;
      }
      const int dn = __builtin_amdgcn_readfirstlane(
          __hip_atomic_load(done, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP));
      if (dn == -2 || (dn >= 0 && t >= dn)) {
        state = dn == -2 ? -1 : 2;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    if (state != 1) {
      ok = state == 2;
      break;
    }
    if (pend) {
      consume();
      pend = false;
    }
    const double* in = ring + (size_t)(t % kPairDepth) * kPairComp * 64;
    pdt = in[8 * 64];
    if (pdt != 0.0) {
#pragma unroll
      for (int c = 0; c < 8; ++c) ps[c] = in[c * 64];
      pst = locate(sv, meta, grid, ps[1], ps[2], ps[3]);
      fetch(sv, pst, pc);
      pend = true;
    }
    publish(consumed, t + 1);  // the slot is in registers now
  }
  if (pend) consume();
  if (valid) img[p] = ok ? I : __builtin_nan("");
}

// ------------------------------------------------------------- triple form
// Three waves per tile on one SIMD -- geodesic, gather (lookup, cell loads,
// interpolation) and emission (fluid scalars, j and alpha, transfer) --
// chained by two LDS rings with the counters of the paired form: three
// independent instruction streams fill the SIMD better than two.  Ring A
// carries (state, dt) from the geodesic wave, ring B the eight interpolated
// primitives and the state in the emission's precision (FP32 values are what
// the FP32 emission would round them to) and dt in FP64 (0: nothing to add).
// Ring depths: ring A (geodesic -> gather) holds 3 rounds in mixed, 2 in
// FP64 (LDS), ring B (gather -> emission) DB_ rounds (3 in mixed where LDS
// allows, else 2).
template <bool MIXED, int DB_>
struct TriRings {
  using T = typename std::conditional<MIXED, float, double>::type;
  static constexpr int kDepthA = MIXED ? 3 : 2, kDepthB = DB_;
  static constexpr int kCompB = 15;                                                  // P[8], s[1..7]
  static constexpr size_t kA = (size_t)kPairTiles * kDepthA * kPairComp * 64;        // doubles
  static constexpr size_t kBd = (size_t)kPairTiles * kDepthB * 64;                   // doubles (dt)
  static constexpr size_t kBt = (size_t)kPairTiles * kDepthB * kCompB * 64;          // elements of T
  static constexpr size_t kLdsExtra = (kA + kBd) * sizeof(double) + kBt * sizeof(T) + kPairTiles * 8 * sizeof(int);
};

template <bool MIXED, int DB_>
constexpr size_t tri_lds_extra() {
  return TriRings<MIXED, DB_>::kLdsExtra;
}

// Polls *cnt until it reaches target: the value read, or -1 after `limit`
// polls.
__device__ __forceinline__ int wait_value(int* cnt, int target, int limit) {
  for (int spin = 0; spin < limit; ++spin) {
    const int v = __builtin_amdgcn_readfirstlane(
        __hip_atomic_load(cnt, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP));
    if (v >= target) return v;
    __builtin_amdgcn_s_sleep(1);
  }
  return -1;
}

// Waits for round t of a ring: 1 ready (*avail = rounds published), 0 the
// writer finished after fewer rounds, -1 failure (the writer gave up, or
// `limit` polls).
__device__ __forceinline__ int wait_round(int* produced, int* done, int t, int* avail, int limit) {
  for (int spin = 0; spin < limit; ++spin) {
    const int pr = __builtin_amdgcn_readfirstlane(
        __hip_atomic_load(produced, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP));
    if (pr > t) {
      *avail = pr;
      return 1;
    }
    const int dn = __builtin_amdgcn_readfirstlane(
        __hip_atomic_load(done, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP));
    if (dn == -2) return -1;
    if (dn >= 0 && t >= dn) return 0;
    __builtin_amdgcn_s_sleep(1);
  }
  return -1;
}

// LDS: the block tables are staged in LDS; otherwise (a snapshot whose tables
// exceed the LDS budget) they are read from global memory, and the gather
// wave keeps each ray's last block record in registers.  ZONE (mixed): 0 the
// zone code is compiled out; 1 the geodesic wave takes FP64 steps while any
// of its rays is in the zone (rp.fp64_r, rp.ring_r), as the other forms do;
// 2 only for launches without a zone, where that test never fires: the zone
// code in the loop layout that measured fastest for launches of several
// workgroup waves (builds 0 and 1 were 5-9 % slower there, same results).
// Build 2's test is per ray, not per tile, so it must never see a zone:
// launch_tri refuses one (MK_ERR_ARG).
template <bool MIXED, bool LDS, int DB_, int ZONE>
__global__ void __launch_bounds__(192 * kPairTiles) k_render_tri(SnapView sv, RenderDev rp, double* __restrict__ img,
                                                                  int32_t* __restrict__ steps, int tab_doubles,
                                                                  int nres) {
  using T = typename std::conditional<MIXED, float, double>::type;
  using R = TriRings<MIXED, DB_>;
  constexpr int DA = R::kDepthA, DB = R::kDepthB, CB = R::kCompB;
  extern __shared__ double lds[];
  double* ringA_all = lds + tab_doubles;
  double* ringBd_all = ringA_all + R::kA;
  T* ringBt_all = reinterpret_cast<T*>(ringBd_all + R::kBd);
  // per tile: producedA, consumedA, doneA, producedB, consumedB, doneB, pad, pad
  int* ctr = reinterpret_cast<int*>(ringBt_all + R::kBt);
  if (threadIdx.x < 8 * kPairTiles) {
    const int f = threadIdx.x & 7;
    ctr[threadIdx.x] = (f == 2 || f == 5) ? -1 : 0;
  }
  const double* meta;
  const int* grid;
  stage_tables<LDS>(sv, lds, meta, grid);  // with LDS: ends with a workgroup barrier
  if constexpr (!LDS) __syncthreads();    // the counters are set
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  // Wave w runs on SIMD w % 4 (the hardware's placement).  The three waves of
  // tile k sit on SIMDs k, k + 1 and k + 2: every SIMD holds one wave of each
  // stage, and a tile that outlives the others of its workgroup has its
  // stages on three SIMDs instead of sharing one.
  const int stage = wave / kPairTiles;
  const int tl = (wave + (kPairTiles - 1) * stage) % kPairTiles;
  double* ringA = ringA_all + (size_t)tl * DA * kPairComp * 64 + lane;
  double* ringBd = ringBd_all + (size_t)tl * DB * 64 + lane;
  T* ringBt = ringBt_all + (size_t)tl * DB * CB * 64 + lane;
  int* c = ctr + 8 * tl;
  const unsigned nwg = gridDim.x, wg = blockIdx.x, per = nwg / 8;
  // with a tile order, workgroups take ranks in index (start) order
  const unsigned lwg = (!rp.order && nwg % 8 == 0) ? (wg % 8) * per + wg / 8 : wg;
  // Workgroups lwg < n (n = nres) are strided: their four tiles come from four
  // quarters of ranks [0, 4n) of the order (four bands of the image without
  // one), taken in snake order (ranks w, 2n - 1 - w, 2n + w, 4n - 1 - w), so
  // the longest tiles share their CU with the shortest.  The host sets n to
  // the whole grid when every workgroup is resident at once, to the resident
  // count when a tile order spreads over at most two waves of workgroups (the
  // first wave then holds one long tile per CU, and the second, the shortest
  // tiles, starts as CUs free), else 0.  Workgroups from n on take the rest of
  // the ranks in order, four adjacent ones each.
  const unsigned n = (unsigned)nres;
  const bool first = lwg < n;
  long tile = first ? ((tl & 1) ? (long)(tl + 1) * n - 1 - lwg : (long)tl * n + lwg)
                    : 4L * n + (long)(lwg - n) * kPairTiles + tl;
  // Strided with a tile order: the three waves of the tile from the first
  // quarter (the longest tiles) issue at a raised priority, so where the waves
  // of a SIMD compete it goes first; the other tiles have slack until it
  // ends, and the waits of every wave sleep, so they still run.  FP64 uses one
  // level per quarter (1-3 % faster there than one raised level; in mixed
  // 1 % slower).
  // s_setprio is a scalar instruction: it takes effect whatever the exec
  // mask, so the condition is made uniform (an SGPR and a scalar branch).
  if (first && rp.order) {
    const int q = __builtin_amdgcn_readfirstlane(tl);
    if constexpr (MIXED) {
      if (q == 0) __builtin_amdgcn_s_setprio(1);
    } else {  // FP64: one level per quarter
      if (q == 0) __builtin_amdgcn_s_setprio(3);
      else if (q == 1) __builtin_amdgcn_s_setprio(2);
      else if (q == 2) __builtin_amdgcn_s_setprio(1);
    }
  }
  if (rp.order && tile < rp.ntiles) tile = rp.order[tile];
  int pi = 0, pj = 0;
  bool valid = tile < rp.ntiles;
  if (valid) {
    pi = (int)(tile / rp.tiles_x) * 8 + (lane >> 3);
    pj = (int)(tile % rp.tiles_x) * 8 + (lane & 7);
    valid = pi < rp.res && pj < rp.res;
  }
  const long p = (long)pi * rp.res + pj;
  if (stage == 0) {  // geodesic
    double s[8], r = 0.0, dt = 0.0, drdl = 1e300;
    int k = 0, nacc = 0;
    bool alive = false;
    if (valid) {
      double sx[4], sv4[4];
      camera_ray(grid_coord(rp.lo, rp.step, pi), grid_coord(rp.lo, rp.step, pj), rp.ci, rp.si, rp.dist, sx, sv4);
      nullify(sx, sv4, rp.a, s);
      r = radius_cal(s[1], s[2], s[3], rp.a);
      dt = step_rule(r, rp.rEH, rp.div, rp.tol);
      alive = true;
    }
    const bool zone = ZONE != 0 && MIXED && (rp.fp64_r > 0.0 || rp.ring_r > 0.0);
    bool ok = true;
    int t = 0, freed = 0;  // freed: ring A rounds the gather wave has read (last value seen)
    while (__any(alive)) {
      // Slot t % DA is free once the gather wave has read round t - DA.  FP64
      // and the zone-free mixed build load that count here without waiting
      // and check it after the RK4 step: a wait at the loop head would also
      // wait on this wave's slot writes of the round before (-2 to -4 % in
      // FP64, -1 % in one-wave mixed launches; the mixed zone build, which
      // also runs launches of several workgroup waves, measured slower).
      int polled = 0;
      if constexpr (!MIXED || ZONE == 0) {
        polled = __hip_atomic_load(c + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      } else if (t - DA + 1 > freed) {
        freed = wait_value(c + 1, t - DA + 1, rp.spin_limit);
        if (freed < 0) {
          ok = false;
          break;
        }
      }
      double* out = ringA + (size_t)(t % DA) * kPairComp * 64;
      if (alive && (k >= rp.N || dt == 0.0)) alive = false;
      // mixed steps the state in place (the step functions write their output
      // only after reading the state; a ray that stops is not used again):
      // -1.3 %; in FP64 the copy is faster
      double nw_[8];
      double* nw = MIXED ? s : nw_;
      if (alive) {
        if constexpr (!MIXED) {
          rk4_step<double>(s, dt, rp.a, nw);
        } else {
          bool fp64;
          if constexpr (ZONE == 2) fp64 = zone && ((r < rp.fp64_r) || (r < rp.ring_r && __builtin_fabs(drdl) < rp.ring_drdl));
          else fp64 = zone && __any((r < rp.fp64_r) || (r < rp.ring_r && __builtin_fabs(drdl) < rp.ring_drdl));
          if (fp64) rk4_step<double>(s, dt, rp.a, nw);
          else rk4_step_mixed(s, dt, (float)rp.a, nw);
        }
      }
      // release the previous round here: its slot writes were issued a whole
      // RK4 step ago, so the release fence does not wait on them
      if (t > 0) publish(c + 0, t);
      if constexpr (!MIXED || ZONE == 0) {
        const int seen = __builtin_amdgcn_readfirstlane(polled);
        if (t - DA + 1 > freed) {
          if (seen >= t - DA + 1) {
            freed = seen;
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
          } else {
            freed = wait_value(c + 1, t - DA + 1, rp.spin_limit);
            if (freed < 0) {
              ok = false;
              break;
            }
          }
        }
      }
      bool wrote = false;
      if (alive) {
        double rn, dtn;
        if constexpr (!MIXED) {
          rn = radius_cal(nw[1], nw[2], nw[3], rp.a);
          dtn = step_rule(rn, rp.rEH, rp.div, rp.tol);
        } else {
          rn = radius_f32(nw[1], nw[2], nw[3], rp.a);
          dtn = step_rule_fast(rn, rp.rEH, rp.inv_div, rp.div, rp.tol);
        }
        if (dtn == 0.0) {
          alive = false;
        } else {
          ++nacc;
          if (k + 1 < rp.N) {  // row N is never emitted
#pragma unroll
            for (int q = 0; q < 8; ++q) out[q * 64] = nw[q];
            out[8 * 64] = dt;
            wrote = true;
          }
          if (zone) drdl = (rn - r) / dt;
          if constexpr (!MIXED) {
#pragma unroll
            for (int q = 0; q < 8; ++q) s[q] = nw[q];
          }
          dt = dtn;
          r = rn;
          if (++k >= rp.N) alive = false;
        }
      }
      if (!wrote) out[8 * 64] = 0.0;
      ++t;
    }
    if (ok) publish(c + 0, t);
    publish(c + 2, ok ? t : -2);
    if (valid && steps) steps[p] = ok ? nacc : -1;
    return;
  }
  if (stage == 1) {  // gather: lookup, cell loads, interpolation
    bool ok = true;
    int t = 0, avail = 0, freed = 0;  // rounds of ring A published, of ring B read (last values seen)
    int cmb = -1;                     // global tables: block of the ray's last sample and its record
    double crec[kMeta];
    for (;; ++t) {
      if (t >= avail) {
        const int st = wait_round(c + 0, c + 2, t, &avail, rp.spin_limit);
        if (st <= 0) {
          ok = st == 0;
          break;
        }
      }
      if (t - DB + 1 > freed) {
        freed = wait_value(c + 4, t - DB + 1, rp.spin_limit);
        if (freed < 0) {
          ok = false;
          break;
        }
      }
      const double* in = ringA + (size_t)(t % DA) * kPairComp * 64;
      double pdt = in[8 * 64];
      double ps[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) ps[q] = in[q * 64];
      publish(c + 1, t + 1);  // ring A slot is in registers
      T* ob = ringBt + (size_t)(t % DB) * CB * 64;
      double* obd = ringBd + (size_t)(t % DB) * 64;
      T P[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) P[q] = T(0);
      if (pdt != 0.0) {
        Stencil stc;
        if constexpr (LDS) stc = locate(sv, meta, grid, ps[1], ps[2], ps[3]);
        else stc = locate_cached(sv, meta, grid, ps[1], ps[2], ps[3], cmb, crec);
        if (stc.mb >= 0) {
          Corners cr;
          fetch(sv, stc, cr);
          interp<T>(stc, cr, P);
        }
        if (stc.mb < 0 || P[0] == T(0)) pdt = 0.0;  // adds exactly nothing
      }
#pragma unroll
      for (int q = 0; q < 8; ++q) ob[q * 64] = P[q];
#pragma unroll
      for (int q = 1; q < 8; ++q) ob[(7 + q) * 64] = (T)ps[q];
      obd[0] = pdt;
      publish(c + 3, t + 1);
    }
    publish(c + 5, ok ? t : -2);
    return;
  }
  // emission and transfer, in step order
  double I = 0.0, trans = 1.0;
  bool ok = true;
  int avail = 0;  // rounds of ring B published (last value seen)
  for (int t = 0;; ++t) {
    if (t >= avail) {
      const int st = wait_round(c + 3, c + 5, t, &avail, rp.spin_limit);
      if (st <= 0) {
        ok = st == 0;
        break;
      }
    }
    const T* ib = ringBt + (size_t)(t % DB) * CB * 64;
    const double pdt = ringBd[(size_t)(t % DB) * 64];
    T P[8];
    double ps[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) P[q] = ib[q * 64];
    ps[0] = 0.0;
#pragma unroll
    for (int q = 1; q < 8; ++q) ps[q] = (double)ib[(7 + q) * 64];
    publish(c + 4, t + 1);  // ring B slot is in registers
    if (pdt != 0.0) {
      T j, al;
      emission_from_prims<T>(rp.ip, P, ps, rp.a, j, al);
      const double cc = -pdt * rp.ip.L_unit;
      I += trans * (cc * (double)j);
      trans *= (1.0 - cc * (double)al);
    }
  }
  if (valid) img[p] = ok ? I : __builtin_nan("");
}

// Critical curve of a distant camera at inclination theta (Bardeen 1973):
// spherical photon orbits of radius r in [r_pro, r_retro] appear at
// alpha = -xi / sin(theta), beta = +-sqrt(eta + a^2 cos^2(theta) -
// xi^2 cot^2(theta)), with xi = (r^2 (3 - r) - a^2 (r + 1)) / (a (r - 1)) and
// eta = r^3 (4 a^2 - r (r - 3)^2) / (a^2 (r - 1)^2).  Image x is alpha and
// image y is beta in this camera's convention.  Returned as a polar table
// R(phi) about the curve's centre (cx, 0), kBins bins; a = 0 gives the circle
// of radius sqrt(27).
namespace {
constexpr int kBins = 256;
void critical_curve_polar(double a, double inc_deg, double* cx, double R[kBins]) {
  const double th = std::min(std::max(inc_deg, 1.0), 179.0) * M_PI / 180.0;
  *cx = 0.0;
  if (std::fabs(a) < 1e-3) {
    for (int k = 0; k < kBins; ++k) R[k] = std::sqrt(27.0);
    return;
  }
  const double aa = std::fabs(a) > 0.999 ? (a > 0 ? 0.999 : -0.999) : a;
  const double r1 = 2.0 * (1.0 + std::cos(2.0 / 3.0 * std::acos(-std::fabs(aa))));
  const double r2 = 2.0 * (1.0 + std::cos(2.0 / 3.0 * std::acos(std::fabs(aa))));
  const double ct = std::cos(th) / std::sin(th), c2 = std::cos(th) * std::cos(th);
  std::vector<double> al, be;
  const int n = 2048;
  for (int q = 0; q <= n; ++q) {
    const double r = std::min(r1, r2) + (std::fabs(r2 - r1)) * q / n;
    if (std::fabs(r - 1.0) < 1e-9) continue;
    const double xi = (r * r * (3.0 - r) - aa * aa * (r + 1.0)) / (aa * (r - 1.0));
    const double eta = r * r * r * (4.0 * aa * aa - r * (r - 3.0) * (r - 3.0)) / (aa * aa * (r - 1.0) * (r - 1.0));
    const double b2 = eta + aa * aa * c2 - xi * xi * ct * ct;
    if (!(b2 >= 0.0)) continue;
    al.push_back(-xi / std::sin(th));
    be.push_back(std::sqrt(b2));
  }
  if (al.size() < 8) {
    for (int k = 0; k < kBins; ++k) R[k] = std::sqrt(27.0);
    return;
  }
  *cx = 0.5 * (*std::min_element(al.begin(), al.end()) + *std::max_element(al.begin(), al.end()));
  std::vector<double> sum(kBins, 0.0);
  std::vector<int> cnt(kBins, 0);
  for (size_t q = 0; q < al.size(); ++q)
    for (int sgn = -1; sgn <= 1; sgn += 2) {
      const double dx = al[q] - *cx, dy = sgn * be[q];
      int k = (int)std::floor((std::atan2(dy, dx) + M_PI) / (2.0 * M_PI) * kBins);
      k = std::min(std::max(k, 0), kBins - 1);
      sum[k] += std::sqrt(dx * dx + dy * dy);
      ++cnt[k];
    }
  // empty bins take the nearest filled bin's radius (circularly)
  for (int k = 0; k < kBins; ++k) {
    if (cnt[k]) {
      R[k] = sum[k] / cnt[k];
      continue;
    }
    for (int d = 1; d < kBins; ++d) {
      const int u = (k + d) % kBins, v = (k - d + kBins) % kBins;
      if (cnt[u]) { R[k] = sum[u] / cnt[u]; break; }
      if (cnt[v]) { R[k] = sum[v] / cnt[v]; break; }
    }
  }
}
}  // namespace

std::vector<int> ring_first_tile_order(int res, double lo, double step, double a, double inc_deg) {
  double cx, R[kBins];
  critical_curve_polar(a, inc_deg, &cx, R);
  const int tx = (res + 7) / 8, nt = tx * tx;
  std::vector<double> d(nt);
  for (int t = 0; t < nt; ++t) {
    // tile centre: pixel 8 t + 3.5 at coordinate (2 i + 1) step + lo
    const double x = (16.0 * (t / tx) + 8.0) * step + lo - cx, y = (16.0 * (t % tx) + 8.0) * step + lo;
    int k = (int)std::floor((std::atan2(y, x) + M_PI) / (2.0 * M_PI) * kBins);
    k = std::min(std::max(k, 0), kBins - 1);
    d[t] = std::fabs(std::sqrt(x * x + y * y) - R[k]);
  }
  std::vector<int> ord(nt);
  for (int t = 0; t < nt; ++t) ord[t] = t;
  std::stable_sort(ord.begin(), ord.end(), [&](int u, int v) { return d[u] < d[v]; });
  return ord;
}

// Device copies of tile orders, one per (device, resolution, fov, spin,
// inclination), made on first use and kept until the process ends (immutable,
// so launches in flight never see them change).  Past kOrderCacheBytes, or
// while the stream is being captured into a graph (no allocation there),
// launches use row-major tiles.
namespace {
std::mutex g_order_mu;
// device table and the host copy it was filled from (kept, so that the
// stream-ordered copy never reads freed memory)
std::map<std::tuple<int, int, double, double, double>, std::pair<int*, std::vector<int>>> g_orders;
size_t g_order_bytes = 0;
constexpr size_t kOrderCacheBytes = (size_t)64 << 20;
}  // namespace

// One tile counter (two ints: next tile, waves done) per (device, stream) for
// the classic form's queue: allocated and zeroed on the stream on first use,
// and left at zero by every launch that uses it.
static std::mutex g_counter_mu;
static std::map<std::pair<int, hipStream_t>, int*> g_counters;

static int* tile_counter(hipStream_t st) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> lk(g_counter_mu);
  auto it = g_counters.find({dev, st});
  if (it != g_counters.end()) return it->second;
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(st, &cap) != hipSuccess || cap != hipStreamCaptureStatusNone) return nullptr;
  int* c = nullptr;
  if (hipMalloc(&c, 2 * sizeof(int)) != hipSuccess) return nullptr;
  if (hipMemsetAsync(c, 0, 2 * sizeof(int), st) != hipSuccess) {
    (void)hipFree(c);
    return nullptr;
  }
  g_counters.emplace(std::make_pair(dev, st), c);
  return c;
}

static const int* render_tile_order(const RenderDev& rp, const mk_render_params* p, hipStream_t st) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  const auto key = std::make_tuple(dev, rp.res, p->fov, p->bhspin, p->inclination);
  std::lock_guard<std::mutex> lk(g_order_mu);
  auto it = g_orders.find(key);
  if (it != g_orders.end()) return it->second.first;
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(st, &cap) != hipSuccess || cap != hipStreamCaptureStatusNone) return nullptr;
  const size_t bytes = (size_t)rp.ntiles * sizeof(int);
  if (g_order_bytes + bytes > kOrderCacheBytes) return nullptr;
  std::vector<int> ord = ring_first_tile_order(rp.res, rp.lo, rp.step, p->bhspin, p->inclination);
  int* d = nullptr;
  if (hipMalloc(&d, bytes) != hipSuccess) return nullptr;
  // ordered on the caller's stream before the launch that reads it
  if (hipMemcpyAsync(d, ord.data(), bytes, hipMemcpyHostToDevice, st) != hipSuccess) {
    (void)hipFree(d);
    return nullptr;
  }
  g_orders.emplace(key, std::make_pair(d, std::move(ord)));
  g_order_bytes += bytes;
  return d;
}

size_t table_bytes(const SnapView& sv) {
  return (size_t)sv.nmb * kMeta * sizeof(double) + (size_t)sv.gx * sv.gy * sv.gz * sizeof(int);
}

int fill_render_dev(const mk_render_params* p, RenderDev* rp) {
  MK_REQUIRE(p->resolution > 0 && p->max_nsteps >= 1 && p->div != 0.0, "render: bad resolution/max_nsteps/div");
  MK_REQUIRE(p->precision == 0 || p->precision == 1, "render: precision must be 0 (plain FP64) or 1 (mixed)");
  MK_REQUIRE(std::fabs(p->bhspin) <= 1.0, "render: |bhspin| must be <= 1");
  MK_REQUIRE(p->observing_frequency > 0.0 && p->M_bh > 0.0 && p->mass_scale > 0.0, "render: bad units");
  const double inc = p->inclination * M_PI / 180;
  rp->a = p->bhspin;
  rp->ci = std::cos(inc);
  rp->si = std::sin(inc);
  rp->dist = p->distance;
  rp->lo = -p->fov / 2.0;
  rp->step = (p->fov / 2.0 - (-p->fov / 2.0)) / (double)(2 * p->resolution);
  rp->rEH = 1.0 + std::sqrt(1.0 - p->bhspin * p->bhspin);
  rp->div = p->div;
  rp->inv_div = 1.0 / p->div;
  rp->tol = p->tol;
  rp->res = p->resolution;
  rp->N = p->max_nsteps;
  rp->tiles_x = (p->resolution + 7) / 8;
  rp->ntiles = (long)rp->tiles_x * rp->tiles_x;
  // grmhd.py:40-55
  const double G = 6.6743e-8, CL = 2.99792458e10, MP = 1.6726e-24, ME = 9.1094e-28;
  const double L = G * p->M_bh / (CL * CL);
  const double dens_unit = p->mass_scale / (L * L * L);
  rp->ip.fluid_gamma = p->fluid_gamma;
  rp->ip.r_high = p->r_high;
  rp->ip.Ne_unit = dens_unit / (MP + ME);
  rp->ip.B_unit = CL * std::sqrt(4.0 * M_PI * dens_unit);
  rp->ip.L_unit = L;
  rp->ip.nu_obs = p->observing_frequency;
  rp->ip.fallback_cos = std::cos(p->fallback_pitch);
  rp->ip.sigma_cut = p->sigma_cut;
  rp->fp64_r = p->fp64_radius;
  rp->ring_r = p->ring_radius;
  rp->ring_drdl = p->ring_drdl;
  rp->spin_limit = (p->flags & MK_RENDER_WAIT_TEST) ? 1 : kSpinLimit;
  rp->order = nullptr;
  rp->next_tile = nullptr;
  return MK_OK;
}

}  // namespace mk

using namespace mk;

// Paired form where the tables fit in LDS next to the rings and the launch
// has fewer than about 1.75 tiles per SIMD (measured on MI355X: paired is
// 15-24 % faster up to 1600 tiles on 1024 SIMDs, and 8-24 % slower from 2304
// tiles, where three classic workgroups per CU keep more waves in flight).
// Tables in LDS when they fit kMaxTableLds and, with the form's rings, the
// 160 KB of a CU; the triple form also runs with the tables in global memory.
template <bool MIXED, bool LDS, int DB_>
static int launch_tri(const SnapView& sv, const RenderDev& rp, double* d_image, int32_t* d_steps, size_t tab,
                      hipStream_t st) {
  long nwg = (rp.ntiles + kPairTiles - 1) / kPairTiles;
  nwg = (nwg + 7) / 8 * 8;
  const size_t lds = tab + tri_lds_extra<MIXED, DB_>();
  // strided tiles for the workgroups resident at once (see k_render_tri)
  // resident workgroups of the build that runs: the zone build (1) for a
  // zone; otherwise the zone-free build (0) when the grid is resident at
  // once with it, else build 2
  const bool zoned = MIXED && (rp.fp64_r > 0.0 || rp.ring_r > 0.0);
  int dev = 0, cus = 0, per_cu = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e == hipSuccess)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, zoned ? k_render_tri<MIXED, LDS, DB_, MIXED ? 1 : 0> : k_render_tri<MIXED, LDS, DB_, 0>,
        192 * kPairTiles, lds);
  if (e == hipSuccess && MIXED && !zoned && nwg > (long)per_cu * cus)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_render_tri<MIXED, LDS, DB_, MIXED ? 2 : 0>,
                                                     192 * kPairTiles, lds);
  if (e != hipSuccess) {
    set_error("mk_render: occupancy query failed");
    return MK_ERR_HIP;
  }
  const long R = (long)per_cu * cus;
  const int nres = nwg <= R ? (int)nwg : (rp.order && nwg <= 2 * R ? (int)R : 0);
  const int tabd = (int)(tab / sizeof(double));
  // mixed: the zone build for a zone; without one, the zone-free build where
  // every workgroup is resident at once (-2 to -3 % at 192^2-256^2) and
  // build 2 in launches of several workgroup waves (see k_render_tri)
  const dim3 grid((unsigned)nwg), blk(192 * kPairTiles);
  if (zoned)
    k_render_tri<MIXED, LDS, DB_, MIXED ? 1 : 0><<<grid, blk, lds, st>>>(sv, rp, d_image, d_steps, tabd, nres);
  else if (MIXED && nwg > R) {
    // build 2 decides FP64 steps per ray: a zone here would break the
    // tile-wide rule every other form follows
    MK_REQUIRE(rp.fp64_r <= 0.0 && rp.ring_r <= 0.0, "mk_render: internal: zone sent to the zone-free build");
    k_render_tri<MIXED, LDS, DB_, MIXED ? 2 : 0><<<grid, blk, lds, st>>>(sv, rp, d_image, d_steps, tabd, nres);
  }
  else
    k_render_tri<MIXED, LDS, DB_, 0><<<grid, blk, lds, st>>>(sv, rp, d_image, d_steps, tabd, nres);
  MK_LAUNCH_CHECK();
  return MK_OK;
}

// The triple form's variant: 0 tables in LDS with 3 ring-B slots (mixed),
// 1 tables in LDS with 2 slots, 2 tables in global memory (3 slots in mixed,
// 2 in FP64).
static int tri_variant(const SnapView& sv, int precision) {
  const size_t tab = (table_bytes(sv) + 15) / 16 * 16;
  const bool small = table_bytes(sv) <= kMaxTableLds;
  if (precision == 1) {
    if (small && tab + tri_lds_extra<true, 3>() <= 160 * 1024) return 0;
    if (small && tab + tri_lds_extra<true, 2>() <= 160 * 1024) return 1;
    return 2;
  }
  return small && tab + tri_lds_extra<false, 2>() <= 160 * 1024 ? 1 : 2;
}

static int choose_form(const SnapView& sv, const RenderDev& rp, int precision, int flags, int* form) {
  const size_t tab = (table_bytes(sv) + 15) / 16 * 16;
  const bool fits2 = table_bytes(sv) <= kMaxTableLds && tab + kPairLdsExtra <= 160 * 1024;
  *form = MK_RENDER_CLASSIC;
  if (flags & MK_RENDER_CLASSIC) return MK_OK;
  if (flags & MK_RENDER_TRIPLE) {
    *form = MK_RENDER_TRIPLE;
    return MK_OK;
  }
  if (flags & MK_RENDER_PAIRED) {
    if (fits2) *form = MK_RENDER_PAIRED;
    return MK_OK;
  }
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
    set_error("mk_render: no device");
    return MK_ERR_HIP;
  }
  // triple form up to 15.5 tiles per CU in mixed precision (a 496^2 image on
  // 256 CUs) and 11 in FP64 (424^2); the classic form beyond, where every
  // SIMD holds three waves of whole rays
  if (2 * rp.ntiles <= (precision == 1 ? 31L : 22L) * cus) *form = MK_RENDER_TRIPLE;
  return MK_OK;
}

extern "C" int mk_render_form(const mk_snapshot* snap, const mk_render_params* p, int32_t* h_form) {
  MK_REQUIRE(snap && p && h_form, "mk_render_form: null argument");
  RenderDev rp;
  int rc = fill_render_dev(p, &rp);
  if (rc) return rc;
  int form = 0;
  rc = choose_form(snap->view(), rp, p->precision, p->flags, &form);
  *h_form = form;
  return rc;
}

extern "C" int mk_render(const mk_snapshot* snap, const mk_render_params* p, double* d_image, int32_t* d_steps,
                         void* stream) {
  MK_REQUIRE(snap && p && d_image, "mk_render: null argument");
  RenderDev rp;
  int rc = fill_render_dev(p, &rp);
  if (rc) return rc;
  const int waves_per_wg = 4;
  long nwg = (rp.ntiles + waves_per_wg - 1) / waves_per_wg;
  nwg = (nwg + 7) / 8 * 8;
  const SnapView sv = snap->view();
  const size_t lds = table_bytes(sv);
  const bool use_lds = lds <= kMaxTableLds;
  const dim3 grid((unsigned)nwg), block(64 * waves_per_wg);
  hipStream_t st = as_stream(stream);
  const size_t tab = (lds + 15) / 16 * 16;
  int form = 0;
  rc = choose_form(sv, rp, p->precision, p->flags, &form);
  if (rc) return rc;
  if (form != MK_RENDER_PAIRED && !(p->flags & MK_RENDER_ROW_MAJOR)) rp.order = render_tile_order(rp, p, st);
  if (form == MK_RENDER_TRIPLE) {
    const int v = tri_variant(sv, p->precision);
    if (p->precision == 0)
      return v == 1 ? launch_tri<false, true, 2>(sv, rp, d_image, d_steps, tab, st)
                    : launch_tri<false, false, 2>(sv, rp, d_image, d_steps, 0, st);
    if (v == 0) return launch_tri<true, true, 3>(sv, rp, d_image, d_steps, tab, st);
    if (v == 1) return launch_tri<true, true, 2>(sv, rp, d_image, d_steps, tab, st);
    return launch_tri<true, false, 3>(sv, rp, d_image, d_steps, 0, st);
  }
  if (form == MK_RENDER_PAIRED) {
    long nwg2 = (rp.ntiles + kPairTiles - 1) / kPairTiles;
    nwg2 = (nwg2 + 7) / 8 * 8;
    const size_t lds2 = tab + kPairLdsExtra;
    const int tabd = (int)(tab / sizeof(double));
    if (p->precision == 0)
      k_render_pair<false><<<dim3((unsigned)nwg2), dim3(128 * kPairTiles), lds2, st>>>(sv, rp, d_image, d_steps, tabd);
    else
      k_render_pair<true><<<dim3((unsigned)nwg2), dim3(128 * kPairTiles), lds2, st>>>(sv, rp, d_image, d_steps, tabd);
    MK_LAUNCH_CHECK();
    return MK_OK;
  }
  // the kernel instance and its LDS: three workgroups per CU (mixed) when
  // their tables fit the CU's 160 KB of LDS
  void (*kern)(SnapView, RenderDev, double*, int32_t*) = nullptr;
  size_t kl = use_lds ? lds : 0;
  if (p->precision == 0) kern = use_lds ? k_render<false, true, false> : k_render<false, false, false>;
  else if (use_lds && 3 * lds <= 160 * 1024) kern = k_render<true, true, true>;
  else kern = use_lds ? k_render<true, true, false> : k_render<true, false, false>;
  dim3 g = grid;
  if (!(p->flags & MK_RENDER_STATIC_TILES)) {  // tiles from a counter (5-12 % faster at 1024^2)
    int dev = 0, cus = 0, per_cu = 0;
    rp.next_tile = tile_counter(st);
    if (rp.next_tile && hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, (int)block.x, kl) == hipSuccess && per_cu > 0) {
      const long resident = (long)per_cu * cus;
      if (resident < nwg) g = dim3((unsigned)resident);
    } else {
      rp.next_tile = nullptr;
    }
  }
  kern<<<g, block, kl, st>>>(sv, rp, d_image, d_steps);
  MK_LAUNCH_CHECK();
  return MK_OK;
}
