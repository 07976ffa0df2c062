// mk_serve.hip -- continuous batching of image and shadow requests.
//
// A worker thread owns one non-blocking stream and iterates:
//   admit   queued requests whose output buffers fit the HBM budget: write
//           their descriptors and a segment (request id, item count) to the
//           device work ring;
//   launch  k_serve: every lane owns a ray slot, runs up to M RK4 steps, and
//           when its image ray or bisection angle finishes it takes the next
//           work item of any admitted request from the ring (bounded atomic
//           retries, so every lane exits);
//   collect per-request done counters; finished outputs are copied to the
//           caller's buffer and their arena blocks freed.
// The launch and the counter copy are one hipGraph replayed per iteration.
// Ray state stays in the slots between launches, so a request's rays see the
// same arithmetic alone or mixed with others; per-ray results equal
// mk_render / mk_shadow_bisect.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <set>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "mk_device.h"
#include "mk_internal.h"
#include "mk_render.h"
#include "mk_snapshot.h"

namespace mk {

enum : int { kKindImage = 1, kKindShadow = 2 };

// Device descriptor of one admitted request.  Whole 256-byte lines: a
// descriptor never shares a cache line with another request's.
struct alignas(256) ReqDev {
  int kind, n_items, res, N, iters, tiled;
  double a, ci, si, dist, lo, step, rEH, div, inv_div, tol;
  ImageParams ip;
  SnapView sv;
  double* out;
  const double* cang;  // shadow: cos / sin of the angles
  const double* sang;
  const int* order;    // image (tiled): the tile taken k-th, longest expected first
};

static_assert(sizeof(ReqDev) % 256 == 0, "descriptor slots are whole cache lines");

// A segment of the work ring: the tiles of one request (64 work items each:
// an 8x8 pixel block, a run of 64 pixels, or 64 bisection angles).
struct Seg {
  int rid, count, next, pad;
};

constexpr int kTile = 64;

// Slot arrays, structure of arrays over L lanes.  f[] holds, per lane:
//   0..7 state, 8 dt, 9 r, 10 I | inner, 11 trans | outer, 12 best,
//   13 r_prev, 14 r_at.   i[]: 0 step, 1 bisection iteration, 2 besti,
//   3 request of the wave, 4 tile of the wave, 5 lane still running.
constexpr int kSlotF = 15, kSlotI = 6;

struct ServeArgs {
  int* done;
  Seg* segs;
  int* seg_head;
  const int* seg_tail;
  int seg_cap;
  double* f;
  int* i;
  unsigned long long* busy;  // lane-steps that advanced a ray (cumulative)
  int L, M;
};

// Wave-level pop of one tile; every lane gets the same (rid, tile), or
// rid = -1 when the ring is drained.  The retry loop is bounded.
__device__ __forceinline__ void pop_tile(const ServeArgs& A, int& rid, int& tile) {
  int g_rid = -1, g_tile = 0;
  if (__lane_id() == 0) {
    const int tail = *A.seg_tail;
    for (int guard = 0; guard < 64; ++guard) {
      const int h = __hip_atomic_load(A.seg_head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (h >= tail) break;
      Seg* sg = A.segs + (h % A.seg_cap);
      const int count = sg->count;
      const int k = atomicAdd(&sg->next, 1);
      if (k < count) {
        g_rid = sg->rid;
        g_tile = k;
        if (k + 1 >= count) atomicCAS(A.seg_head, h, h + 1);
        break;
      }
      atomicCAS(A.seg_head, h, h + 1);
    }
  }
  rid = __builtin_amdgcn_readfirstlane(__shfl(g_rid, 0));
  tile = __builtin_amdgcn_readfirstlane(__shfl(g_tile, 0));
}

struct Lane {
  double s[8], dt, r, f10, f11, best, r_prev, r_at;
  int k, n, besti;
  int cmb;              // block of the ray's last sample (-1: none), and its record:
  double crec[kMeta];   // the snapshot tables are in global memory here, not in LDS
};

__device__ __forceinline__ void shadow_begin(const ReqDev& q, int item, Lane& L) {
  const double mid = (L.f11 - L.f10) / 2 + L.f10;  // (outer - inner)/2 + inner
  double sx[4], sv4[4];
  camera_ray(q.cang[item] * mid, q.sang[item] * mid, q.ci, q.si, q.dist, sx, sv4);
  nullify(sx, sv4, q.a, L.s);
  L.r = radius_cal(L.s[1], L.s[2], L.s[3], q.a);
  L.dt = step_rule(L.r, q.rEH, 40.0, 1e-2);
  L.k = 0;
  L.best = -__builtin_inf();
  L.besti = -1;
  L.r_prev = __builtin_nan("");
  L.r_at = __builtin_nan("");
}

// Image item -> pixel index i*res+j.  Items run over 8x8 tiles when the
// resolution allows, so a wave's 64 items are neighbouring rays.
__device__ __forceinline__ int pixel_of_item(const ReqDev& q, int item) {
  if (!q.tiled) return item;
  const int tile = q.order[item >> 6], w = item & 63, tx = q.res >> 3;
  const int i = (tile / tx) * 8 + (w >> 3), j = (tile % tx) * 8 + (w & 7);
  return i * q.res + j;
}

// Starts work item `item` of request q in lane state L.  Returns true when
// the item is already finished (a bisection with no halvings keeps the
// initial inner radius 0.5).
__device__ __forceinline__ bool item_begin(const ReqDev& q, int item, Lane& L) {
  if (q.kind == kKindImage) {
    const int px = pixel_of_item(q, item);
    const int i = px / q.res, j = px % q.res;
    double sx[4], sv4[4];
    camera_ray(grid_coord(q.lo, q.step, i), grid_coord(q.lo, q.step, j), q.ci, q.si, q.dist, sx, sv4);
    nullify(sx, sv4, q.a, L.s);
    L.r = radius_cal(L.s[1], L.s[2], L.s[3], q.a);
    L.dt = step_rule(L.r, q.rEH, q.div, q.tol);
    L.f10 = 0.0;  // I
    L.f11 = 1.0;  // transmission
    L.k = 0;
    L.n = 0;
    L.cmb = -1;
    return false;
  }
  L.f10 = 0.5;   // inner
  L.f11 = 10.0;  // outer
  L.n = 0;       // bisection iteration
  if (q.iters == 0) return true;
  shadow_begin(q, item, L);
  return false;
}

// A sample fetched one step ahead (its emission is added after the next
// RK4 step, so the cell loads overlap that step's arithmetic).
struct Pending {
  Stencil st;
  Corners c;
  double s[8], dt;
};

template <bool MIXED>
__device__ __forceinline__ void consume(const ReqDev& q, const Pending& p, Lane& L) {
  using T = typename std::conditional<MIXED, float, double>::type;
  if (p.st.mb < 0) return;
  T P[8], j, al;
  interp<T>(p.st, p.c, P);
  if (P[0] == T(0)) return;
  emission_from_prims<T>(q.ip, P, p.s, q.a, j, al);
  const double cc = -p.dt * q.ip.L_unit;
  L.f10 += L.f11 * (cc * (double)j);
  L.f11 *= (1.0 - cc * (double)al);
}

// One pass of the image-ray loop (the body of the fused render loop, split
// around the single place where a pending sample is consumed).  `go` false
// only settles the pending sample (the wave is out of steps).  Returns true
// when the ray is finished.  Keeping one call site for consume() keeps the
// arithmetic of every sample the same wherever a launch boundary falls.
// IN_PLACE: the sample is consumed right after its fetch instead of after
// the next step (the same arithmetic on the same values, so the same bits),
// so no corner records are held across a step: fewer registers.
template <bool MIXED, bool IN_PLACE = false>
__device__ __forceinline__ bool image_pass(const ReqDev& q, bool go, Lane& L, Pending& p, bool& pend) {
  bool fin = false, adv = false;
  double nw[8];
  if (go) {
    if (L.k >= q.N || L.dt == 0.0) {
      fin = true;
    } else {
      if constexpr (!MIXED) rk4_step<double>(L.s, L.dt, q.a, nw);
      else rk4_step_mixed(L.s, L.dt, (float)q.a, nw);
      adv = true;
    }
  }
  if (pend) {
    consume<MIXED>(q, p, L);
    pend = false;
  }
  if (adv) {
    double rn, dtn;
    if constexpr (!MIXED) {
      rn = radius_cal(nw[1], nw[2], nw[3], q.a);
      dtn = step_rule(rn, q.rEH, q.div, q.tol);
    } else {
      rn = radius_f32(nw[1], nw[2], nw[3], q.a);
      dtn = step_rule_fast(rn, q.rEH, q.inv_div, q.div, q.tol);
    }
    if (dtn == 0.0) {
      fin = true;
    } else {
      if (L.k + 1 < q.N) {  // row N is never emitted
        // FP64 has no registers to spare for the cached record (it would spill)
        if constexpr (MIXED) p.st = locate_cached(q.sv, q.sv.meta, q.sv.grid, nw[1], nw[2], nw[3], L.cmb, L.crec);
        else p.st = locate(q.sv, q.sv.meta, q.sv.grid, nw[1], nw[2], nw[3]);
        fetch(q.sv, p.st, p.c);
#pragma unroll
        for (int c = 0; c < 8; ++c) p.s[c] = nw[c];
        p.dt = L.dt;
        pend = true;
        if constexpr (IN_PLACE) {
          consume<MIXED>(q, p, L);
          pend = false;
        }
      }
#pragma unroll
      for (int c = 0; c < 8; ++c) L.s[c] = nw[c];
      L.dt = dtn;
      L.r = rn;
      ++L.k;
      fin = L.k >= q.N;  // no sample is pending then
    }
  }
  return fin;
}

// One row of the bisection's integration (the body of walk_ray), and the
// bisection update when an integration ends.  Returns true when the angle's
// last iteration is done (its radius is in L.f10).
__device__ __forceinline__ bool shadow_step(const ReqDev& q, int item, Lane& L) {
  const double r_here = L.r;
  bool frozen = (L.dt == 0.0);
  double nw[8], rn = 0.0, dtn = 0.0;
  if (!frozen) {
    rk4_step<double>(L.s, L.dt, q.a, nw);
    rn = radius_cal(nw[1], nw[2], nw[3], q.a);
    dtn = step_rule(rn, q.rEH, 40.0, 1e-2);
    frozen = (dtn == 0.0);
  }
  const double row_dt = frozen ? 0.0 : L.dt;
  if (row_dt > L.best) {
    L.best = row_dt;
    L.besti = L.k;
    L.r_at = L.r_prev;
  }
  bool end = frozen;
  if (!end) {
    L.r_prev = r_here;
    if (L.k == q.N - 1) {
      end = true;
    } else {
#pragma unroll
      for (int c = 0; c < 8; ++c) L.s[c] = nw[c];
      L.dt = dtn;
      L.r = rn;
      ++L.k;
    }
  }
  if (!end) return false;
  const double sel = (L.besti >= 1) ? L.r_at : L.r;
  const double mid = (L.f11 - L.f10) / 2 + L.f10;
  if (sel < 100.0) L.f10 = mid;
  else L.f11 = mid;
  if (++L.n < q.iters) {
    shadow_begin(q, item, L);
    return false;
  }
  return true;
}

// Advance kernel.  Each wave owns one tile (64 work items of one request)
// at a time: its lanes step their rays together, as in the render kernel,
// so neighbouring rays share cache lines; when every lane of the tile is
// done the wave takes the next tile of any admitted request.  A wave stops
// after M steps and leaves its rays in the slot arrays for the next launch.
template <bool MIXED>
__global__ void __launch_bounds__(256, 2) k_serve(const ReqDev* __restrict__ req, ServeArgs A) {
  const int l = blockIdx.x * blockDim.x + threadIdx.x;  // L is a multiple of 256
  const int lane = threadIdx.x & 63;
  const long L_ = A.L;
  int rid = __builtin_amdgcn_readfirstlane(A.i[3 * L_ + l]);
  int tile = __builtin_amdgcn_readfirstlane(A.i[4 * L_ + l]);
  Lane L{};
  L.cmb = -1;
  bool live = false;
  if (rid >= 0) {
    live = A.i[5 * L_ + l] != 0;
#pragma unroll
    for (int c = 0; c < 8; ++c) L.s[c] = A.f[c * L_ + l];
    L.dt = A.f[8 * L_ + l];
    L.r = A.f[9 * L_ + l];
    L.f10 = A.f[10 * L_ + l];
    L.f11 = A.f[11 * L_ + l];
    L.best = A.f[12 * L_ + l];
    L.r_prev = A.f[13 * L_ + l];
    L.r_at = A.f[14 * L_ + l];
    L.k = A.i[l];
    L.n = A.i[L_ + l];
    L.besti = A.i[2 * L_ + l];
  }
  int budget = A.M, steps = 0;
  while (budget > 0) {
    if (rid < 0) {
      pop_tile(A, rid, tile);
      if (rid < 0) break;  // ring drained
      const ReqDev& q = req[rid];
      const int item = tile * kTile + lane;
      live = item < q.n_items;
      if (live && item_begin(q, item, L)) {
        q.out[item] = L.f10;
        live = false;
      }
    }
    const ReqDev& q = req[rid];
    const int item = tile * kTile + lane;
    if (q.kind == kKindImage) {
      Pending p;
      bool pend = false;
      while (true) {
        const bool go = budget > 0 && __any(live);  // uniform
        if (live) {
          steps += go;
          if (image_pass<MIXED>(q, go, L, p, pend)) {
            q.out[pixel_of_item(q, item)] = L.f10;
            live = false;
          }
        }
        if (!go) break;  // the last pass only settled pending samples
        --budget;
      }
    } else {
      while (budget > 0 && __any(live)) {
        if (live) {
          ++steps;
          if (shadow_step(q, item, L)) {
            q.out[item] = L.f10;
            live = false;
          }
        }
        --budget;
      }
    }
    if (!__any(live)) {  // tile done: count its items once
      if (lane == 0) atomicAdd(&A.done[rid], min(kTile, q.n_items - tile * kTile));
      rid = -1;
    }
  }
  unsigned long long w = (unsigned long long)steps;
  for (int off = 32; off > 0; off >>= 1) w += __shfl_xor(w, off);
  if (lane == 0) atomicAdd(A.busy, w);
  A.i[3 * L_ + l] = rid;
  A.i[4 * L_ + l] = tile;
  A.i[5 * L_ + l] = live ? 1 : 0;
  if (rid >= 0) {
#pragma unroll
    for (int c = 0; c < 8; ++c) A.f[c * L_ + l] = L.s[c];
    A.f[8 * L_ + l] = L.dt;
    A.f[9 * L_ + l] = L.r;
    A.f[10 * L_ + l] = L.f10;
    A.f[11 * L_ + l] = L.f11;
    A.f[12 * L_ + l] = L.best;
    A.f[13 * L_ + l] = L.r_prev;
    A.f[14 * L_ + l] = L.r_at;
    A.i[l] = L.k;
    A.i[L_ + l] = L.n;
    A.i[2 * L_ + l] = L.besti;
  }
}

// ------------------------------------------------------------ persistent form
// One launch serves requests until the host asks it to stop: work admitted
// while it runs starts without a launch boundary.  Tiles (64 work items) are
// numbered by ticket in the order the host publishes them.  A wave takes the
// next ticket with one fetch-and-add, published or not, waits until the host
// has published it, runs the tile to the end and takes the next ticket.
// (Claiming only published tickets needs a compare-and-swap loop; when work
// appears, a thousand waiting waves retrying it cost about N^2/2 atomic round
// trips, and the launch served a twentieth of the relaunched form's rate.)
// While the launch runs the host issues no copy to or from device memory: it
// publishes work in pinned, coherent host memory (segments, descriptors, tile
// orders and angle tables) and the waves read it there; finished pixels and
// radii go straight to pinned host memory, and the wave that completes a
// request stores the request's sequence number in a pinned completion flag.
// Every wait is bounded: a waiting wave sleeps between polls and exits when
// the host sets `stop`, or reports an error after `idle_limit` sleeps.  The
// host sets `stop` only when every published tile is done, and publishes
// nothing more until the launch has ended, so the waves leave holding
// unpublished tickets and the next launch resumes the count at the published
// total.  Descriptor slots and input tables are fresh memory for every
// request of a launch (the host relaunches before it reuses any), so no cache
// of a running launch can hold an older request's copy of them.

struct PSeg {          // one published request: tickets [start, start + count)
  long long start;
  int rid, count, seq, desc;
};

struct PCtl {          // pinned host memory; each word on its own 128-byte line
  int stop, err, pad0[30];
  long long pub_tiles;
  long long pad1[15];
  int pub_segs, pad2[31];
};

struct PArgs {
  PCtl* ctl;                    // host
  const PSeg* segs;             // host, seg_cap entries
  int* fin;                     // host, per request slot
  unsigned long long* taken;    // device: tickets handed out
  unsigned long long* seen;     // device: largest published count a wave has read, | kHalt once stopping
  int* done;                    // device: finished items per request slot
  unsigned long long* busy;     // device: lane-steps that advanced a ray
  unsigned long long* total;    // device: 64 x the wave-steps of every tile run
  int seg_cap, idle_limit;
};

constexpr unsigned long long kHalt = 1ull << 62;  // in `seen`: the host has set stop
constexpr long long kFrontTickets = 64;           // waiters this close to the published count poll host memory

__device__ __forceinline__ long long bcast64(long long v) {
  const int lo = __builtin_amdgcn_readfirstlane(__shfl((int)(v & 0xffffffffLL), 0));
  const int hi = __builtin_amdgcn_readfirstlane(__shfl((int)(v >> 32), 0));
  return (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

// Synchronisation without acquires.  On this GPU an acquire at agent or
// system scope invalidates the L2 of the wave's XCD, and the busy waves of that
// XCD would fetch their snapshot blocks from HBM again; the relaunched form's
// loop has no acquire either.  Everything a wave reads that the host or
// another wave wrote during the launch is read where no cache holds it:
// segments, published counts, tile orders and angle tables are in pinned
// coherent host memory, which the L2 does not cache, and are read with
// system-scope loads; descriptors are read through the scalar cache, which the
// launch started empty, from slots written before their segment was published
// and never rewritten during the launch; the device counters are accessed with
// agent-scope atomics, which are coherent across XCDs.  Program order orders a
// wave's reads: each read is issued after the value it depends on has returned
// (the segment after the published count, the descriptor after the segment).

// Lane 0 waits until ticket t is published: 1 when it is, 0 when the host
// has set stop, -1 after `idle_limit` sleeps.  The device word `seen` is read
// every round; waiters within kFrontTickets of it also read host memory every
// round (so new work is seen within one round), the others every 64 rounds.
// A waiter that reads stop in host memory marks `seen` for all the others.
__device__ __forceinline__ int wait_published(const PArgs& P, long long t) {
  int state = 0;
  if (__lane_id() == 0) {
    for (int round = 0, slept = 0;; ++round) {
      const unsigned long long v = __hip_atomic_load(P.seen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      long long pub = (long long)(v & ~kHalt);
      if (t < pub) { state = 1; break; }
      if (v & kHalt) break;
      const long long d = t - pub;
      if (d < kFrontTickets || ((round + (int)t) & 63) == 0) {
        pub = __hip_atomic_load(&P.ctl->pub_tiles, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __hip_atomic_fetch_max(P.seen, (unsigned long long)pub, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (t < pub) { state = 1; break; }
        if (__hip_atomic_load(&P.ctl->stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM)) {
          __hip_atomic_fetch_or(P.seen, kHalt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
      }
      // about 13 us a round near the published count, 54 us far from it
      const int naps = d < 512 ? 4 : 16;
      for (int k = 0; k < naps; ++k) __builtin_amdgcn_s_sleep(127);
      slept += naps;
      if (slept >= P.idle_limit) { state = -1; break; }
    }
  }
  return __builtin_amdgcn_readfirstlane(__shfl(state, 0));
}

// Lane 0 finds the published segment holding ticket t (a binary search over
// the last seg_cap segments, from the wave's cursor); false if none does.
__device__ __forceinline__ bool find_segment(const PArgs& P, long long t, int& cursor, PSeg& out) {
  int ok = 0;
  PSeg e{};
  if (__lane_id() == 0) {
    const int nseg = __hip_atomic_load(&P.ctl->pub_segs, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    int lo = cursor > nseg - P.seg_cap ? cursor : nseg - P.seg_cap, hi = nseg - 1;
    if (lo < 0) lo = 0;
    if (lo < hi) {  // most tickets lie in the wave's current segment or the next one
      const PSeg* m = P.segs + (lo % P.seg_cap);
      const int sq = __hip_atomic_load(&m->seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      const long long st = __hip_atomic_load(&m->start, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      const int ct = __hip_atomic_load(&m->count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      if (sq == lo && st <= t && t < st + ct) hi = lo;
    }
    // entries older than the ones in flight may be overwritten by newer
    // ones (seq differs); they lie before t's segment
    for (int guard = 0; guard < 40 && lo < hi; ++guard) {
      const int mid = lo + (hi - lo + 1) / 2;
      const PSeg* m = P.segs + (mid % P.seg_cap);
      const int sq = __hip_atomic_load(&m->seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      const long long st = __hip_atomic_load(&m->start, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      if (sq != mid || st <= t) lo = mid;
      else hi = mid - 1;
    }
    if (lo >= 0 && lo < nseg) {
      const PSeg* m = P.segs + (lo % P.seg_cap);
      e.seq = __hip_atomic_load(&m->seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      e.start = __hip_atomic_load(&m->start, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      e.count = __hip_atomic_load(&m->count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      e.rid = __hip_atomic_load(&m->rid, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      e.desc = __hip_atomic_load(&m->desc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      ok = (e.seq == lo && e.start <= t && t < e.start + e.count) ? 1 : 0;
      if (ok) cursor = lo;
    }
  }
  ok = __builtin_amdgcn_readfirstlane(__shfl(ok, 0));
  out.start = bcast64(e.start);
  out.rid = __builtin_amdgcn_readfirstlane(__shfl(e.rid, 0));
  out.count = __builtin_amdgcn_readfirstlane(__shfl(e.count, 0));
  out.seq = __builtin_amdgcn_readfirstlane(__shfl(e.seq, 0));
  out.desc = __builtin_amdgcn_readfirstlane(__shfl(e.desc, 0));
  return ok != 0;
}

// `req` is the descriptor slots in pinned host memory, read as the relaunched
// form reads its table: scalar loads through a read-only pointer.  SHADOW:
// the launch serves bisections too; without them the mixed image path fits
// three waves per SIMD (168 VGPRs against 236: the bisection's FP64 walk
// needs the rest), and the host publishes no bisection to such a launch.
template <bool MIXED, bool SHADOW>
__global__ void __launch_bounds__(256, (MIXED && !SHADOW) ? 3 : 2)
    k_serve_persistent(const ReqDev* __restrict__ req, PArgs P) {
  const int lane = threadIdx.x & 63;
  int cursor = 0;
  while (true) {
    unsigned long long tk = 0;
    if (lane == 0) tk = __hip_atomic_fetch_add(P.taken, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const long long t = bcast64((long long)tk);
    const int state = wait_published(P, t);
    if (state <= 0) {
      if (state < 0 && lane == 0) __hip_atomic_store(&P.ctl->err, 3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      break;
    }
    PSeg sg;
    if (!find_segment(P, t, cursor, sg)) {
      if (lane == 0) __hip_atomic_store(&P.ctl->err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      break;
    }
    const ReqDev& q = req[sg.desc];
    const int tile = (int)(t - sg.start), rid = sg.rid;
    const int item = tile * kTile + lane;
    unsigned steps = 0;   // this tile's lane-steps, added to the busy count at its end
    unsigned wsteps = 0;  // and its wave-steps
    Lane L{};
    L.cmb = -1;
    bool live = item < q.n_items;
    if (live && item_begin(q, item, L)) {
      q.out[item] = L.f10;
      live = false;
    }
    if (q.kind == kKindImage) {
      // the tile order lives in host memory: read it once per tile
      const int px = live ? pixel_of_item(q, item) : 0;
      Pending p;
      bool pend = false;
      while (true) {
        const bool go = __any(live);
        wsteps += go;
        if (live) {
          steps += go;
          if (image_pass<MIXED, true>(q, go, L, p, pend)) {
            q.out[px] = L.f10;
            live = false;
          }
        }
        if (!go) break;
      }
    } else if constexpr (SHADOW) {
      while (__any(live)) {
        ++wsteps;
        if (live) {
          ++steps;
          if (shadow_step(q, item, L)) {
            q.out[item] = L.f10;
            live = false;
          }
        }
      }
    } else if (lane == 0) {  // the host publishes no bisection to this launch
      __hip_atomic_store(&P.ctl->err, 4, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
    unsigned long long w = steps;
    for (int off = 32; off > 0; off >>= 1) w += __shfl_xor(w, off);
    // The tile's outputs are complete in host memory (a system-scope release:
    // write-back and wait, no invalidation) before its items are counted, so
    // when the request's last tile sees the full count every output of the
    // request is already there, and that wave's flag store follows it.
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
    if (lane == 0) {
      atomicAdd(P.busy, w);
      atomicAdd(P.total, 64ull * wsteps);
      const int n = min(kTile, q.n_items - tile * kTile);
      const int old = __hip_atomic_fetch_add(&P.done[rid], n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (old + n >= q.n_items) {
        __hip_atomic_store(&P.done[rid], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
        __hip_atomic_store(&P.fin[rid], sg.seq + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      }
    }
  }
}

__global__ void k_set_pair(unsigned long long* p, unsigned long long a, unsigned long long b) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    p[0] = a;
    p[1] = b;
  }
}

__global__ void k_fill_int(int* p, long n, int v) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

}  // namespace mk

using namespace mk;
using Clock = std::chrono::steady_clock;

struct mk_server {
  mk_server_config cfg{};
  hipStream_t stream = nullptr;
  int dev = 0, L = 0, M = 0, Rmax = 0, seg_cap = 4096;
  // device
  ReqDev* d_req = nullptr;
  int* d_done = nullptr;
  Seg* d_segs = nullptr;
  int* d_head = nullptr;
  int* d_tail = nullptr;
  double* d_f = nullptr;
  int* d_i = nullptr;
  char* d_arena = nullptr;
  size_t arena_bytes = 0;
  // pinned host mirrors
  ReqDev* h_req = nullptr;
  Seg* h_segs = nullptr;
  int* h_done = nullptr;   // Rmax done counters, then head, tail, busy (u64)
  int* h_tail = nullptr;
  int* h_zero = nullptr;
  hipGraphExec_t graph = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // host state
  std::vector<const mk_snapshot*> snaps;
  std::map<size_t, size_t> free_blocks;  // offset -> size
  size_t in_use = 0;
  struct Pending {
    int64_t id;
    int kind;
    int priority;
    mk_image_request img;
    mk_shadow_request sh;
    std::vector<double> angles;
    double* h_out;
    size_t bytes;
    Clock::time_point t0;
  };
  struct Flight {
    int64_t id;
    int n_items;
    size_t off, bytes, out_bytes;
    double* h_out;
    Clock::time_point t0;
    std::vector<double> cs;  // shadow: cos, sin of the angles (copy source)
    std::vector<int> ord;    // image: tile order (copy source)
    int seq = -1;            // persistent form: the request's segment number
    bool shadow = false;     // persistent form: a bisection
  };
  // persistent form (MK_SERVER_PERSISTENT): pinned, coherent host memory
  // shared with the running launch, and its device counters
  bool persistent = false;
  PCtl* p_ctl = nullptr;
  PSeg* p_segs = nullptr;
  ReqDev* p_desc = nullptr;
  int* p_fin = nullptr;
  char* p_in = nullptr;              // tile orders and angle tables of a launch
  size_t p_in_cap = 0, p_in_off = 0;
  char* p_out = nullptr;             // output staging (the budgeted arena)
  unsigned long long* d_p = nullptr; // tickets taken, published count seen, busy and total lane-steps
  int* d_pdone = nullptr;
  double* d_direct = nullptr;        // staging of bisections the idle worker runs itself
  double* d_direct_img = nullptr;    // and of its images
  long long pub_tiles = 0;
  long long resume_ticket = 0;       // the next launch's first ticket: the published count at the last stop
  int seg_seq = 0, desc_seq = 0, desc_cap = 0;
  bool launched = false, draining = false, stopping = false;
  bool default_lanes = true;         // lanes not set by the caller: two waves per SIMD
  bool launch_shadow = false;        // the running launch serves bisections too
  bool launch_deep = false;          // the next launch starts with a deep backlog
  bool want_shadow = false;          // a bisection waits for a launch that serves it
  std::chrono::steady_clock::time_point launch_t0;
  std::map<std::pair<int, double>, std::vector<int>> orders;  // tile orders by (resolution, fov)
  std::deque<Pending> waiting;         // by priority, then submit order
  Pending running{};                   // the request the idle worker runs itself
  int64_t running_id = 0;              // its id while it runs, else 0
  std::set<int64_t> cancelled;         // ids removed from the queue (the newest kCancelledKept)
  std::deque<int64_t> cancelled_order;
  std::map<int, Flight> flight;
  std::vector<int> free_rids;
  std::vector<std::pair<int64_t, double>> unpolled;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::thread worker;
  std::atomic<bool> stop{false};
  int64_t next_id = 1;
  int tail = 0, head = 0;
  int status = MK_OK;
  std::string err;
  mk_server_stats st{};
};

namespace {

size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

constexpr size_t kMaxUnpolled = size_t(1) << 20;  // completions kept for mk_server_poll
constexpr int kDirectShadowAngles = 1024;         // bisections an idle server runs itself
constexpr int kDirectImageRes = 512;              // and images, when one is all it has
constexpr size_t kCancelledKept = size_t(1) << 16; // cancelled ids remembered for mk_server_wait

// Queues a request behind every queued one of the same or a higher priority
// (caller holds mu).
void enqueue(mk_server* s, mk_server::Pending&& p) {
  auto it = s->waiting.end();
  while (it != s->waiting.begin() && std::prev(it)->priority < p.priority) --it;
  s->waiting.insert(it, std::move(p));
}

bool arena_alloc(mk_server* s, size_t bytes, size_t* off) {
  for (auto it = s->free_blocks.begin(); it != s->free_blocks.end(); ++it) {
    if (it->second >= bytes) {
      *off = it->first;
      const size_t rest = it->second - bytes;
      const size_t o = it->first;
      s->free_blocks.erase(it);
      if (rest) s->free_blocks[o + bytes] = rest;
      s->in_use += bytes;
      return true;
    }
  }
  return false;
}

void arena_free(mk_server* s, size_t off, size_t bytes) {
  s->in_use -= bytes;
  auto it = s->free_blocks.emplace(off, bytes).first;
  auto nx = std::next(it);
  if (nx != s->free_blocks.end() && it->first + it->second == nx->first) {
    it->second += nx->second;
    s->free_blocks.erase(nx);
  }
  if (it != s->free_blocks.begin()) {
    auto pv = std::prev(it);
    if (pv->first + pv->second == it->first) {
      pv->second += it->second;
      s->free_blocks.erase(it);
    }
  }
}

// A failed HIP call sets the server's error and returns false: SRV_HIP where
// the caller holds mu (or no other thread runs yet), SRV_HIP_UNLOCKED where it
// does not (the worker's waits on the GPU run without mu, so that submit and
// stats calls are not blocked).
#define SRV_HIP(call)                                                      \
  do {                                                                     \
    hipError_t e_ = (call);                                                \
    if (e_ != hipSuccess) {                                                \
      s->status = MK_ERR_HIP;                                              \
      s->err = std::string(#call) + ": " + hipGetErrorString(e_);          \
      return false;                                                        \
    }                                                                      \
  } while (0)
#define SRV_HIP_UNLOCKED(call)                                             \
  do {                                                                     \
    hipError_t e_ = (call);                                                \
    if (e_ != hipSuccess) {                                                \
      std::lock_guard<std::mutex> lk_(s->mu);                              \
      s->status = MK_ERR_HIP;                                              \
      s->err = std::string(#call) + ": " + hipGetErrorString(e_);          \
      return false;                                                        \
    }                                                                      \
  } while (0)

size_t image_out_bytes(int res) { return align_up((size_t)res * res * sizeof(double)); }

size_t image_bytes(int res) {
  const size_t tiles = (res % 8 == 0) ? (size_t)(res / 8) * (res / 8) : 0;
  return image_out_bytes(res) + align_up(tiles * sizeof(int));
}

// An image's tile order: the Schwarzschild critical circle (radius sqrt(27) M)
// for every request, cached per (resolution, fov).  The spin's own curve
// orders a little better but costs the admitting thread about 0.3 ms per
// 512^2 request, which the latency of a Poisson stream shows (p50 8.9 ->
// 9.2 ms).
std::vector<int> tile_order(mk_server* s, int res, double fov, double lo, double step) {
  const auto key = std::make_pair(res, fov);
  auto it = s->orders.find(key);
  if (it != s->orders.end()) return it->second;
  std::vector<int> ord = ring_first_tile_order(res, lo, step, 0.0, 90.0);
  if (s->orders.size() >= 256) s->orders.clear();
  s->orders.emplace(key, ord);
  return ord;
}

// mk_render's parameters of an image request (make_image's fixed div, tol,
// pitch and sigma cut).
mk_render_params image_params(const mk_server* s, const mk_image_request& r) {
  mk_render_params rpar{};
  rpar.bhspin = r.bhspin;
  rpar.inclination = r.inclination;
  rpar.distance = r.distance;
  rpar.fov = r.fov;
  rpar.mass_scale = r.mass_scale;
  rpar.M_bh = r.M_bh;
  rpar.r_high = r.r_high;
  rpar.fluid_gamma = r.fluid_gamma;
  rpar.observing_frequency = r.observing_frequency;
  rpar.div = 40.0;
  rpar.tol = 1e-4;
  rpar.fallback_pitch = M_PI / 3.0;
  rpar.sigma_cut = 100.0;
  rpar.resolution = r.resolution;
  rpar.max_nsteps = r.max_nsteps;
  rpar.precision = s->cfg.precision;
  return rpar;
}

// Descriptor fields of a request, all but its memory (out, order, cang,
// sang); cs receives a bisection's cos and sin table (cos then sin).
void fill_desc(mk_server* s, const mk_server::Pending& p, ReqDev& q, std::vector<double>* cs) {
  std::memset(&q, 0, sizeof q);
  if (p.kind == kKindImage) {
    const mk_image_request& r = p.img;
    const mk_render_params rpar = image_params(s, r);
    RenderDev rd;
    fill_render_dev(&rpar, &rd);
    q.kind = kKindImage;
    q.n_items = r.resolution * r.resolution;
    q.res = r.resolution;
    q.tiled = (r.resolution % 8 == 0) ? 1 : 0;
    q.N = r.max_nsteps;
    q.a = rd.a; q.ci = rd.ci; q.si = rd.si; q.dist = rd.dist; q.lo = rd.lo; q.step = rd.step;
    q.rEH = rd.rEH; q.div = rd.div; q.inv_div = rd.inv_div; q.tol = rd.tol;
    q.ip = rd.ip;
    q.sv = s->snaps[r.snapshot]->view();
    return;
  }
  const mk_shadow_request& r = p.sh;
  const int n = r.n_angles;
  int iters = 0;
  double inner = 0.5, outer = 10.0;
  while (outer - inner > r.error_allowed && iters < r.max_it) {
    inner = (outer - inner) / 2 + inner;
    ++iters;
  }
  const double inc = r.inclination * M_PI / 180;
  q.kind = kKindShadow;
  q.n_items = n;
  q.N = r.max_steps;
  q.iters = iters;
  q.a = r.bhspin;
  q.ci = std::cos(inc);
  q.si = std::sin(inc);
  q.dist = r.distance;
  q.rEH = 1.0 + std::sqrt(1.0 - r.bhspin * r.bhspin);
  cs->resize(2 * (size_t)n);
  for (int k = 0; k < n; ++k) {
    (*cs)[k] = std::cos(p.angles[k]);
    (*cs)[n + k] = std::sin(p.angles[k]);
  }
}

// Moves waiting requests into flight while they fit (caller holds mu).
bool admit(mk_server* s) {
  while (!s->waiting.empty() && !s->free_rids.empty() && s->tail - s->head < s->seg_cap) {
    mk_server::Pending& p = s->waiting.front();
    size_t off;
    if (!arena_alloc(s, p.bytes, &off)) break;  // back-pressure: wait for memory
    const int rid = s->free_rids.back();
    s->free_rids.pop_back();
    ReqDev& q = s->h_req[rid];
    mk_server::Flight f{p.id, 0, off, p.bytes, 0, p.h_out, p.t0, {}, {}};
    fill_desc(s, p, q, &f.cs);
    q.out = (double*)(s->d_arena + off);
    if (q.kind == kKindImage) {
      if (q.tiled) {
        f.ord = tile_order(s, p.img.resolution, p.img.fov, q.lo, q.step);
        q.order = (const int*)(s->d_arena + off + image_out_bytes(p.img.resolution));
      }
    } else {
      q.cang = (double*)(s->d_arena + off) + q.n_items;
      q.sang = q.cang + q.n_items;
    }
    f.n_items = q.n_items;
    f.out_bytes = (size_t)q.n_items * sizeof(double);
    SRV_HIP(hipMemcpyAsync(s->d_req + rid, &q, sizeof q, hipMemcpyHostToDevice, s->stream));
    SRV_HIP(hipMemcpyAsync(s->d_done + rid, s->h_zero, sizeof(int), hipMemcpyHostToDevice, s->stream));
    Seg& sg = s->h_segs[s->tail % s->seg_cap];
    sg.rid = rid;
    sg.count = (q.n_items + kTile - 1) / kTile;
    sg.next = 0;
    sg.pad = 0;
    SRV_HIP(hipMemcpyAsync(s->d_segs + (s->tail % s->seg_cap), &sg, sizeof sg, hipMemcpyHostToDevice, s->stream));
    ++s->tail;
    s->flight[rid] = std::move(f);
    if (q.kind == kKindImage && q.tiled) {
      const std::vector<int>& ord = s->flight[rid].ord;
      SRV_HIP(hipMemcpyAsync((void*)q.order, ord.data(), ord.size() * sizeof(int), hipMemcpyHostToDevice, s->stream));
    }
    if (q.kind == kKindShadow) {
      // the Flight entry owns the host table until the request completes
      const std::vector<double>& cs = s->flight[rid].cs;
      SRV_HIP(hipMemcpyAsync((void*)q.cang, cs.data(), cs.size() * sizeof(double), hipMemcpyHostToDevice,
                             s->stream));
    }
    s->waiting.pop_front();
    ++s->st.admitted;
  }
  *s->h_tail = s->tail;
  SRV_HIP(hipMemcpyAsync(s->d_tail, s->h_tail, sizeof(int), hipMemcpyHostToDevice, s->stream));
  return true;
}

ServeArgs make_args(mk_server* s) {
  ServeArgs A;
  A.done = s->d_done;
  A.segs = s->d_segs;
  A.seg_head = s->d_head;
  A.seg_tail = s->d_tail;
  A.seg_cap = s->seg_cap;
  A.f = s->d_f;
  A.i = s->d_i;
  A.busy = (unsigned long long*)(s->d_head + 2);
  A.L = s->L;
  A.M = s->M;
  return A;
}

bool enqueue_iteration(mk_server* s) {
  const ServeArgs A = make_args(s);
  const unsigned blocks = (unsigned)((s->L + 255) / 256);
  if (s->cfg.precision == 0) k_serve<false><<<blocks, 256, 0, s->stream>>>(s->d_req, A);
  else k_serve<true><<<blocks, 256, 0, s->stream>>>(s->d_req, A);
  SRV_HIP_UNLOCKED(hipGetLastError());
  SRV_HIP_UNLOCKED(hipMemcpyAsync(s->h_done, s->d_done, s->Rmax * sizeof(int), hipMemcpyDeviceToHost, s->stream));
  SRV_HIP_UNLOCKED(hipMemcpyAsync(s->h_done + s->Rmax, s->d_head, 4 * sizeof(int), hipMemcpyDeviceToHost, s->stream));
  return true;
}

bool build_graph(mk_server* s) {
  hipGraph_t g;
  SRV_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
  const bool ok = enqueue_iteration(s);
  hipError_t e = hipStreamEndCapture(s->stream, &g);
  if (!ok) return false;
  SRV_HIP(e);
  SRV_HIP(hipGraphInstantiate(&s->graph, g, nullptr, nullptr, 0));
  SRV_HIP(hipGraphDestroy(g));
  return true;
}

bool iterate(mk_server* s) {
  SRV_HIP_UNLOCKED(hipEventRecord(s->ev0, s->stream));
  if (s->graph) {
    SRV_HIP_UNLOCKED(hipGraphLaunch(s->graph, s->stream));
  } else if (!enqueue_iteration(s)) {
    return false;
  }
  SRV_HIP_UNLOCKED(hipEventRecord(s->ev1, s->stream));
  SRV_HIP_UNLOCKED(hipEventSynchronize(s->ev1));
  float ms = 0;
  SRV_HIP_UNLOCKED(hipEventElapsedTime(&ms, s->ev0, s->ev1));
  std::lock_guard<std::mutex> lk(s->mu);
  s->st.gpu_ms += ms;
  ++s->st.iterations;
  s->st.lane_steps_total += (int64_t)s->L * s->M;
  return true;
}

// Finished requests: copy out, free, report (caller holds mu).
bool collect(mk_server* s) {
  s->head = s->h_done[s->Rmax];
  unsigned long long busy;
  std::memcpy(&busy, s->h_done + s->Rmax + 2, sizeof busy);
  s->st.lane_steps_busy = (int64_t)busy;
  std::vector<int> fin;
  for (auto& kv : s->flight)
    if (s->h_done[kv.first] >= kv.second.n_items) fin.push_back(kv.first);
  for (int rid : fin) {
    const mk_server::Flight& f = s->flight[rid];
    SRV_HIP(hipMemcpyAsync(f.h_out, s->d_arena + f.off, f.out_bytes, hipMemcpyDeviceToHost, s->stream));
    SRV_HIP(hipStreamSynchronize(s->stream));
    const int64_t id = f.id;
    const int n_items = f.n_items;
    const double ms = std::chrono::duration<double, std::milli>(Clock::now() - f.t0).count();
    arena_free(s, f.off, f.bytes);
    s->flight.erase(rid);
    s->free_rids.push_back(rid);
    if (s->unpolled.size() >= kMaxUnpolled)  // keep the newest completions
      s->unpolled.erase(s->unpolled.begin(), s->unpolled.begin() + kMaxUnpolled / 2);
    s->unpolled.emplace_back(id, ms);
    ++s->st.completed;
    s->st.items_done += n_items;
  }
  if (!fin.empty()) s->cv_done.notify_all();
  return true;
}

// An idle server (no launch running, nothing in flight) runs a small request
// at the head of its queue itself: a bisection of up to 1024 angles in
// mk_shadow_bisect's band schedule (about one integration chain where the
// server's launches take one per halving), and, when it is the only request
// queued, an image of up to 512^2 through mk_render (three waves per 8x8 tile
// up to 256^2, with row-major tiles, which at one tile per SIMD cost nothing;
// above, mk_render's ring-first tile order, built once per view and kept).
// The same results, bit for bit.  (Caller holds mu.)
bool direct_next(mk_server* s) {
  if (!s->flight.empty() || s->waiting.empty()) return false;
  const mk_server::Pending& p = s->waiting.front();
  if (p.kind != kKindImage)
    return !(s->cfg.flags & MK_SERVER_SHADOW_IN_LAUNCH) && p.sh.n_angles <= kDirectShadowAngles;
  return !(s->cfg.flags & MK_SERVER_IMAGE_IN_LAUNCH) && s->waiting.size() == 1 &&
         p.img.resolution <= kDirectImageRes;
}

// Runs the request at the head of the queue, which direct_next chose: it
// moves to the running slot, where a wait on it still finds it, and mu is
// released while the GPU works.  Only the worker touches the slot.
bool serve_direct(mk_server* s) {
  const mk_snapshot* snap = nullptr;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->running = std::move(s->waiting.front());
    s->waiting.pop_front();
    s->running_id = s->running.id;
    if (s->running.kind == kKindImage) snap = s->snaps[s->running.img.snapshot];
  }
  const mk_server::Pending* p = &s->running;
  long items = 0;
  SRV_HIP_UNLOCKED(hipEventRecord(s->ev0, s->stream));
  if (p->kind == kKindImage) {
    mk_render_params rp = image_params(s, p->img);
    if (rp.resolution <= 256) rp.flags |= MK_RENDER_ROW_MAJOR;  // one tile per SIMD: order costs nothing
    items = (long)rp.resolution * rp.resolution;
    if (mk_render(snap, &rp, s->d_direct_img, nullptr, s->stream) != MK_OK) {
      std::lock_guard<std::mutex> lk(s->mu);
      s->status = MK_ERR_HIP;
      s->err = std::string("direct image: ") + mk_last_error();
      return false;
    }
    SRV_HIP_UNLOCKED(hipMemcpyAsync(p->h_out, s->d_direct_img, items * sizeof(double), hipMemcpyDeviceToHost,
                                    s->stream));
  } else {
    const mk_shadow_request& r = p->sh;
    items = r.n_angles;
    SRV_HIP_UNLOCKED(shadow_bisect_on_stream(r.bhspin, r.inclination, p->angles.data(), r.n_angles, r.max_steps,
                                             r.error_allowed, r.max_it, r.distance, s->d_direct, p->h_out,
                                             s->stream));
  }
  SRV_HIP_UNLOCKED(hipEventRecord(s->ev1, s->stream));
  SRV_HIP_UNLOCKED(hipEventSynchronize(s->ev1));
  float gpu_ms = 0;
  SRV_HIP_UNLOCKED(hipEventElapsedTime(&gpu_ms, s->ev0, s->ev1));
  std::lock_guard<std::mutex> lk(s->mu);
  const double ms = std::chrono::duration<double, std::milli>(Clock::now() - p->t0).count();
  if (s->unpolled.size() >= kMaxUnpolled)
    s->unpolled.erase(s->unpolled.begin(), s->unpolled.begin() + kMaxUnpolled / 2);
  s->unpolled.emplace_back(p->id, ms);
  s->st.gpu_ms += gpu_ms;
  ++s->st.admitted;
  ++s->st.completed;
  s->st.items_done += items;
  s->running_id = 0;
  s->cv_done.notify_all();
  return true;
}

void worker_main(mk_server* s) {
  if (hipSetDevice(s->dev) != hipSuccess) {
    std::lock_guard<std::mutex> lk(s->mu);
    s->status = MK_ERR_HIP;
    s->err = "worker: hipSetDevice failed";
    s->cv_done.notify_all();
    return;
  }
  while (!s->stop.load()) {
    {
      std::unique_lock<std::mutex> lk(s->mu);
      if (direct_next(s)) {
        lk.unlock();
        if (!serve_direct(s)) break;
        continue;
      }
      if (!admit(s)) break;
      if (s->flight.empty()) {
        s->cv_work.wait_for(lk, std::chrono::milliseconds(20));
        continue;
      }
    }
    if (!iterate(s)) break;
    std::lock_guard<std::mutex> lk(s->mu);
    if (!collect(s)) break;
  }
  std::lock_guard<std::mutex> lk(s->mu);
  s->cv_done.notify_all();
}

// ---------------------------------------------------------- persistent form
// The host ends a launch once nothing has been queued or in flight for
// kPersistIdleMs, or, when a launch older than kPersistMaxAgeMs finds nothing
// in flight, at once: a waiting wave then holds a ticket for at most that long
// plus the longest request.  A wave reports an error after kPersistWaitSleeps
// sleeps of about 3.4 us (about 28 s).
constexpr int kPersistIdleMs = 20;
// A mixed launch without bisections runs three waves per SIMD when the work
// queued or in flight is at least this many 64-item tiles per wave of the
// default two-waves-per-SIMD launch.
constexpr int kDeepTilesPerWave = 4;
// A launch about to start without such a backlog first waits up to this long
// for more work (a batch being submitted); only idle-to-busy starts pay it.
constexpr int kPersistGatherUs = 200;
constexpr int kPersistMaxAgeMs = 1000;
constexpr int kPersistWaitSleeps = 1 << 23;
constexpr size_t kPersistInputBytes = (size_t)256 << 20;  // tile orders and angle tables per launch
constexpr int kPersistDescSlots = 16384;          // descriptor slots per launch

template <typename T>
void store_shared(T* p, T v) {
  __atomic_store_n(p, v, __ATOMIC_RELEASE);
}
template <typename T>
T load_shared(const T* p) {
  return __atomic_load_n(p, __ATOMIC_ACQUIRE);
}

// Publishes waiting requests to the running (or next) launch while they fit
// (caller holds mu).  A request whose tables or descriptor would not fit the
// launch's fresh memory starts a drain: nothing more is admitted until every
// request in flight has finished and the launch has ended.
void admit_persistent(mk_server* s) {
  if (s->launched && s->stopping) return;  // the waves may be leaving: nothing is published until the launch ends
  bool any = false;
  while (!s->draining && !s->waiting.empty() && !s->free_rids.empty()) {
    mk_server::Pending& p = s->waiting.front();
    if (p.kind != kKindImage && s->launched && !s->launch_shadow) {
      // a launch without the bisection path: hold the queue here (FIFO)
      // until that launch has drained and ended
      s->want_shadow = true;
      break;
    }
    size_t in_bytes = 0;
    if (p.kind == kKindImage) {
      const int res = p.img.resolution;
      in_bytes = (res % 8 == 0) ? align_up((size_t)(res / 8) * (res / 8) * sizeof(int)) : 0;
    } else {
      in_bytes = align_up((size_t)2 * p.sh.n_angles * sizeof(double));
    }
    if (s->p_in_off + in_bytes > s->p_in_cap || s->desc_seq >= s->desc_cap) {
      s->draining = true;
      break;
    }
    const size_t out_bytes = align_up((size_t)(p.kind == kKindImage ? p.img.resolution * p.img.resolution
                                                                     : p.sh.n_angles) * sizeof(double));
    size_t off;
    if (!arena_alloc(s, out_bytes, &off)) break;  // back-pressure
    const int rid = s->free_rids.back();
    s->free_rids.pop_back();
    mk_server::Flight f{p.id, 0, off, out_bytes, 0, p.h_out, p.t0, {}, {}};
    f.shadow = p.kind != kKindImage;
    ReqDev q;
    fill_desc(s, p, q, &f.cs);
    q.out = (double*)(s->p_out + off);
    char* in = s->p_in + s->p_in_off;
    if (q.kind == kKindImage) {
      if (q.tiled) {
        const std::vector<int>& ord = tile_order(s, p.img.resolution, p.img.fov, q.lo, q.step);
        std::memcpy(in, ord.data(), ord.size() * sizeof(int));
        q.order = (const int*)in;
      }
    } else {
      std::memcpy(in, f.cs.data(), f.cs.size() * sizeof(double));
      q.cang = (const double*)in;
      q.sang = q.cang + q.n_items;
      f.cs.clear();
    }
    s->p_in_off += in_bytes;
    const int slot = s->desc_seq++;
    std::memcpy(&s->p_desc[slot], &q, sizeof q);
    f.n_items = q.n_items;
    f.out_bytes = (size_t)q.n_items * sizeof(double);
    f.seq = s->seg_seq;
    store_shared(&s->p_fin[rid], 0);
    PSeg& e = s->p_segs[s->seg_seq % s->seg_cap];
    e.start = s->pub_tiles;
    e.rid = rid;
    e.count = (q.n_items + kTile - 1) / kTile;
    e.desc = slot;
    store_shared(&e.seq, s->seg_seq);
    s->pub_tiles += e.count;
    ++s->seg_seq;
    s->flight[rid] = std::move(f);
    s->waiting.pop_front();
    ++s->st.admitted;
    any = true;
  }
  if (any) {  // the segment count first: a wave that sees a ticket sees its segment
    store_shared(&s->p_ctl->pub_segs, s->seg_seq);
    store_shared(&s->p_ctl->pub_tiles, s->pub_tiles);
  }
}

// Finished requests: their completion flag holds seq + 1 (caller holds mu).
void collect_persistent(mk_server* s) {
  std::vector<int> fin;
  for (auto& kv : s->flight)
    if (load_shared(&s->p_fin[kv.first]) == kv.second.seq + 1) fin.push_back(kv.first);
  for (int rid : fin) {
    const mk_server::Flight& f = s->flight[rid];
    std::memcpy(f.h_out, s->p_out + f.off, f.out_bytes);
    const double ms = std::chrono::duration<double, std::milli>(Clock::now() - f.t0).count();
    const int64_t id = f.id;
    const int n_items = f.n_items;
    arena_free(s, f.off, f.bytes);
    s->flight.erase(rid);
    s->free_rids.push_back(rid);
    if (s->unpolled.size() >= kMaxUnpolled)
      s->unpolled.erase(s->unpolled.begin(), s->unpolled.begin() + kMaxUnpolled / 2);
    s->unpolled.emplace_back(id, ms);
    ++s->st.completed;
    s->st.items_done += n_items;
  }
  if (!fin.empty()) s->cv_done.notify_all();
}

bool launch_persistent(mk_server* s) {
  PArgs P;
  P.ctl = s->p_ctl;
  P.segs = s->p_segs;
  P.fin = s->p_fin;
  P.taken = s->d_p;
  P.seen = s->d_p + 1;
  P.busy = s->d_p + 2;
  P.total = s->d_p + 3;
  P.done = s->d_pdone;
  P.seg_cap = s->seg_cap;
  P.idle_limit = kPersistWaitSleeps;
  store_shared(&s->p_ctl->stop, 0);
  s->stopping = false;
  s->want_shadow = false;
  bool shadow = false;
  for (const auto& kv : s->flight) shadow = shadow || kv.second.shadow;
  s->launch_shadow = shadow;
  // mixed without bisections, default lanes and a deep backlog: three waves
  // per SIMD, for throughput; otherwise two, each wave faster, for latency
  unsigned blocks = (unsigned)((s->L + 255) / 256);
  if (s->cfg.precision == 1 && !shadow && s->default_lanes && s->launch_deep) blocks = blocks * 3 / 2;
  // tickets below resume_ticket are done; the previous launch's waves left
  // holding tickets from it on, unpublished then
  k_set_pair<<<1, 64, 0, s->stream>>>(s->d_p, (unsigned long long)s->resume_ticket,
                                      (unsigned long long)s->resume_ticket);
  SRV_HIP_UNLOCKED(hipGetLastError());
  SRV_HIP_UNLOCKED(hipEventRecord(s->ev0, s->stream));
  if (s->cfg.precision == 0) {
    if (shadow) k_serve_persistent<false, true><<<blocks, 256, 0, s->stream>>>(s->p_desc, P);
    else k_serve_persistent<false, false><<<blocks, 256, 0, s->stream>>>(s->p_desc, P);
  } else {
    if (shadow) k_serve_persistent<true, true><<<blocks, 256, 0, s->stream>>>(s->p_desc, P);
    else k_serve_persistent<true, false><<<blocks, 256, 0, s->stream>>>(s->p_desc, P);
  }
  SRV_HIP_UNLOCKED(hipGetLastError());
  SRV_HIP_UNLOCKED(hipEventRecord(s->ev1, s->stream));
  s->launched = true;
  s->launch_t0 = Clock::now();
  std::lock_guard<std::mutex> lk(s->mu);
  ++s->st.iterations;
  return true;
}

// Asks the running launch to end; called only when every published tile is
// done (caller holds mu).  Nothing is published until the launch has ended.
void stop_persistent(mk_server* s) {
  if (s->stopping) return;
  s->stopping = true;
  s->resume_ticket = s->pub_tiles;
  store_shared(&s->p_ctl->stop, 1);
}

// The launch has ended: its time and lane-steps go to the statistics.
bool end_persistent(mk_server* s) {
  float ms = 0;
  SRV_HIP_UNLOCKED(hipEventElapsedTime(&ms, s->ev0, s->ev1));
  unsigned long long c[2] = {0, 0};  // busy and total lane-steps, over every launch so far
  SRV_HIP_UNLOCKED(hipMemcpy(c, s->d_p + 2, sizeof c, hipMemcpyDeviceToHost));
  std::lock_guard<std::mutex> lk(s->mu);
  s->st.gpu_ms += ms;
  s->st.lane_steps_busy = (int64_t)c[0];
  s->st.lane_steps_total = (int64_t)c[1];
  s->launched = false;
  return true;
}

void worker_persistent(mk_server* s) {
  if (hipSetDevice(s->dev) != hipSuccess) {
    std::lock_guard<std::mutex> lk(s->mu);
    s->status = MK_ERR_HIP;
    s->err = "worker: hipSetDevice failed";
    s->cv_done.notify_all();
    return;
  }
  bool ok = true;
  Clock::time_point last_busy = Clock::now(), ready_since{};
  bool ready = false;  // work admitted, no launch yet (since ready_since)
  while (ok && !s->stop.load()) {
    if (s->launched) {
      const hipError_t q = hipEventQuery(s->ev1);
      if (q == hipSuccess) ok = end_persistent(s);
      else if (q != hipErrorNotReady) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->status = MK_ERR_HIP;
        s->err = std::string("persistent launch: ") + hipGetErrorString(q);
        break;
      }
    }
    std::unique_lock<std::mutex> lk(s->mu);
    if (const int ec = load_shared(&s->p_ctl->err)) {
      s->status = MK_ERR_STATE;
      s->err = ec == 1   ? "persistent launch: a ticket without a published segment"
               : ec == 4 ? "persistent launch: a bisection reached a launch without the bisection path"
                         : "persistent launch: a wave waited longer than its limit for work";
      break;
    }
    collect_persistent(s);
    if (!s->launched && direct_next(s)) {
      lk.unlock();
      ok = serve_direct(s);
      last_busy = Clock::now();
      continue;
    }
    if (s->draining && s->flight.empty()) {
      if (s->launched) {
        stop_persistent(s);  // the launch ends within about a tenth of a millisecond
      } else {  // the next launch starts with fresh descriptor slots and tables
        s->p_in_off = 0;
        s->desc_seq = 0;
        s->draining = false;
      }
    }
    admit_persistent(s);
    const Clock::time_point now = Clock::now();
    if (!s->flight.empty() || !s->waiting.empty()) {
      last_busy = now;
    }
    if (s->launched && s->flight.empty() &&
        (s->want_shadow || (s->waiting.empty() && now - last_busy > std::chrono::milliseconds(kPersistIdleMs)) ||
         now - s->launch_t0 > std::chrono::milliseconds(kPersistMaxAgeMs))) {
      stop_persistent(s);  // nothing in flight: the next request relaunches
    }
    if (!s->launched && !s->flight.empty()) {
      long long items = 0;  // queued or in flight, counted here under mu
      for (const auto& kv : s->flight) items += kv.second.n_items;
      for (const auto& p : s->waiting)
        items += p.kind == kKindImage ? (long long)p.img.resolution * p.img.resolution : p.sh.n_angles;
      s->launch_deep = items >= (long long)kDeepTilesPerWave * s->L;
      if (!ready) {
        ready = true;
        ready_since = now;
      }
      if (s->launch_deep || now - ready_since >= std::chrono::microseconds(kPersistGatherUs)) {
        ready = false;
        lk.unlock();
        ok = launch_persistent(s);
        last_busy = Clock::now();
        continue;
      }
    }
    if (s->flight.empty() && s->waiting.empty() && !s->launched)
      s->cv_work.wait_for(lk, std::chrono::milliseconds(20));
    else
      s->cv_work.wait_for(lk, std::chrono::microseconds(50));
  }
  store_shared(&s->p_ctl->stop, 1);  // the waves finish every published tile, then leave
  (void)hipStreamSynchronize(s->stream);
  std::lock_guard<std::mutex> lk(s->mu);
  s->cv_done.notify_all();
}

int server_error(mk_server* s) {
  if (s->status == MK_OK) return MK_OK;
  set_error("mk_server: %s", s->err.c_str());
  return s->status;
}

}  // namespace

extern "C" {

int mk_server_create(const mk_server_config* cfg, mk_server** out) {
  MK_REQUIRE(cfg && out, "mk_server_create: null argument");
  MK_REQUIRE(cfg->precision == 0 || cfg->precision == 1, "mk_server_create: precision must be 0 or 1");
  *out = nullptr;
  mk_server* s = new mk_server();
  s->cfg = *cfg;
  int dev = 0, cus = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  s->dev = dev;
  s->default_lanes = cfg->lanes <= 0;
  s->L = cfg->lanes > 0 ? cfg->lanes : std::max(cus, 1) * 4 * 2 * 64;
  s->L = (s->L + 255) / 256 * 256;  // whole blocks: every wave is full
  s->M = cfg->steps_per_launch > 0 ? cfg->steps_per_launch : 256;
  s->Rmax = cfg->max_requests > 0 ? cfg->max_requests : 1024;
  s->persistent = (cfg->flags & MK_SERVER_PERSISTENT) != 0;
  // the persistent form stages outputs in pinned host memory: 1 GiB by default
  s->arena_bytes = cfg->hbm_budget > 0 ? cfg->hbm_budget : ((size_t)(s->persistent ? 1 : 4) << 30);
  const long L = s->L;
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc(&s->d_direct, 3 * kDirectShadowAngles * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&s->d_direct_img, (size_t)kDirectImageRes * kDirectImageRes * sizeof(double));
  if (s->persistent) {
    // seg_cap > Rmax: a segment's entry is overwritten only long after its
    // request finished (at most Rmax requests are in flight)
    s->seg_cap = std::max(4096, 4 * s->Rmax);
    s->desc_cap = kPersistDescSlots;
    s->p_in_cap = kPersistInputBytes;
    const unsigned fl = hipHostMallocCoherent | hipHostMallocMapped;
    if (e == hipSuccess) e = hipHostMalloc((void**)&s->p_ctl, sizeof(PCtl), fl);
    if (e == hipSuccess) e = hipHostMalloc((void**)&s->p_segs, s->seg_cap * sizeof(PSeg), fl);
    if (e == hipSuccess) e = hipHostMalloc((void**)&s->p_desc, s->desc_cap * sizeof(ReqDev), fl);
    if (e == hipSuccess) e = hipHostMalloc((void**)&s->p_fin, s->Rmax * sizeof(int), fl);
    if (e == hipSuccess) e = hipHostMalloc((void**)&s->p_in, s->p_in_cap, fl);
    if (e == hipSuccess) e = hipHostMalloc((void**)&s->p_out, s->arena_bytes, fl);
    if (e == hipSuccess) e = hipMalloc(&s->d_p, 4 * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMalloc(&s->d_pdone, s->Rmax * sizeof(int));
    // the launch dereferences these host pointers: they must be its addresses too
    for (void* h : {(void*)s->p_ctl, (void*)s->p_segs, (void*)s->p_desc, (void*)s->p_fin, (void*)s->p_in,
                    (void*)s->p_out}) {
      void* d = nullptr;
      if (e == hipSuccess) e = hipHostGetDevicePointer(&d, h, 0);
      if (e == hipSuccess && d != h) e = hipErrorInvalidValue;
    }
    if (e == hipSuccess) {
      std::memset(s->p_ctl, 0, sizeof(PCtl));
      std::memset(s->p_segs, 0xff, s->seg_cap * sizeof(PSeg));  // seq -1: no segment yet
      std::memset(s->p_fin, 0, s->Rmax * sizeof(int));
      k_fill_int<<<1, 64, 0, s->stream>>>((int*)s->d_p, 8, 0);
      k_fill_int<<<grid_for(s->Rmax, 256), 256, 0, s->stream>>>(s->d_pdone, s->Rmax, 0);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventCreate(&s->ev0);
    if (e == hipSuccess) e = hipEventCreate(&s->ev1);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) {
      set_error("mk_server_create: %s", hipGetErrorString(e));
      mk_server_destroy(s);
      return e == hipErrorOutOfMemory ? MK_ERR_NOMEM : MK_ERR_HIP;
    }
    s->free_blocks[0] = s->arena_bytes;
    for (int r = s->Rmax - 1; r >= 0; --r) s->free_rids.push_back(r);
    s->st.hbm_budget = s->arena_bytes;
    s->worker = std::thread(worker_persistent, s);
    *out = s;
    return MK_OK;
  }
  if (e == hipSuccess) e = hipMalloc(&s->d_req, s->Rmax * sizeof(ReqDev));
  if (e == hipSuccess) e = hipMalloc(&s->d_done, s->Rmax * sizeof(int));
  if (e == hipSuccess) e = hipMalloc(&s->d_segs, s->seg_cap * sizeof(Seg));
  if (e == hipSuccess) e = hipMalloc(&s->d_head, 4 * sizeof(int));
  if (e == hipSuccess) s->d_tail = s->d_head + 1;
  if (e == hipSuccess) e = hipMalloc(&s->d_f, (size_t)kSlotF * L * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&s->d_i, (size_t)kSlotI * L * sizeof(int));
  if (e == hipSuccess) e = hipMalloc(&s->d_arena, s->arena_bytes);
  if (e == hipSuccess) e = hipHostMalloc(&s->h_req, s->Rmax * sizeof(ReqDev), hipHostMallocDefault);
  if (e == hipSuccess) e = hipHostMalloc(&s->h_segs, s->seg_cap * sizeof(Seg), hipHostMallocDefault);
  if (e == hipSuccess) e = hipHostMalloc(&s->h_done, (s->Rmax + 4) * sizeof(int), hipHostMallocDefault);
  if (e == hipSuccess) e = hipHostMalloc(&s->h_tail, 2 * sizeof(int), hipHostMallocDefault);
  if (e == hipSuccess) {
    s->h_zero = s->h_tail + 1;
    *s->h_zero = 0;
    *s->h_tail = 0;
  }
  if (e == hipSuccess) e = hipEventCreate(&s->ev0);
  if (e == hipSuccess) e = hipEventCreate(&s->ev1);
  if (e == hipSuccess) {
    // slots empty (rid = -1), counters zero: fill kernels, no memset nodes
    k_fill_int<<<grid_for(kSlotI * L, 256), 256, 0, s->stream>>>(s->d_i, kSlotI * L, -1);
    k_fill_int<<<grid_for(s->Rmax, 256), 256, 0, s->stream>>>(s->d_done, s->Rmax, 0);
    k_fill_int<<<1, 64, 0, s->stream>>>(s->d_head, 4, 0);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  if (e != hipSuccess) {
    set_error("mk_server_create: %s", hipGetErrorString(e));
    mk_server_destroy(s);
    return e == hipErrorOutOfMemory ? MK_ERR_NOMEM : MK_ERR_HIP;
  }
  s->free_blocks[0] = s->arena_bytes;
  for (int r = s->Rmax - 1; r >= 0; --r) s->free_rids.push_back(r);
  s->st.hbm_budget = s->arena_bytes;
  if (cfg->use_graph && !build_graph(s)) {
    set_error("mk_server_create: graph capture failed: %s", s->err.c_str());
    mk_server_destroy(s);
    return MK_ERR_HIP;
  }
  s->worker = std::thread(worker_main, s);
  *out = s;
  return MK_OK;
}

int mk_server_add_snapshot(mk_server* s, const mk_snapshot* snap, int32_t* id) {
  MK_REQUIRE(s && snap && id, "mk_server_add_snapshot: null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  s->snaps.push_back(snap);
  *id = (int32_t)s->snaps.size() - 1;
  return MK_OK;
}

int mk_server_submit_image(mk_server* s, const mk_image_request* r, double* h_out, int64_t* id) {
  MK_REQUIRE(s && r && h_out && id, "mk_server_submit_image: null argument");
  MK_REQUIRE(r->resolution > 0 && r->resolution <= 8192 && r->max_nsteps >= 1, "mk_server_submit_image: bad size");
  MK_REQUIRE(std::fabs(r->bhspin) <= 1.0 && r->observing_frequency > 0 && r->M_bh > 0 && r->mass_scale > 0,
             "mk_server_submit_image: bad physical parameters");
  std::lock_guard<std::mutex> lk(s->mu);
  if (int rc = server_error(s)) return rc;
  MK_REQUIRE(r->snapshot >= 0 && r->snapshot < (int)s->snaps.size(), "mk_server_submit_image: unknown snapshot");
  mk_server::Pending p{};
  p.id = s->next_id++;
  p.kind = kKindImage;
  p.priority = r->priority;
  p.img = *r;
  p.h_out = h_out;
  p.bytes = image_bytes(r->resolution);
  if (p.bytes > s->arena_bytes) {
    set_error("mk_server_submit_image: %zu bytes of output exceed the %zu-byte budget", p.bytes, s->arena_bytes);
    return MK_ERR_NOMEM;
  }
  p.t0 = Clock::now();
  *id = p.id;
  enqueue(s, std::move(p));
  ++s->st.submitted;
  s->cv_work.notify_one();
  return MK_OK;
}

int mk_server_submit_shadow(mk_server* s, const mk_shadow_request* r, double* h_out, int64_t* id) {
  MK_REQUIRE(s && r && h_out && id && r->h_angles, "mk_server_submit_shadow: null argument");
  MK_REQUIRE(r->n_angles > 0 && r->max_steps >= 1 && std::fabs(r->bhspin) <= 1.0, "mk_server_submit_shadow: bad size");
  std::lock_guard<std::mutex> lk(s->mu);
  if (int rc = server_error(s)) return rc;
  mk_server::Pending p{};
  p.id = s->next_id++;
  p.kind = kKindShadow;
  p.priority = r->priority;
  p.sh = *r;
  p.angles.assign(r->h_angles, r->h_angles + r->n_angles);
  p.sh.h_angles = nullptr;
  p.h_out = h_out;
  p.bytes = align_up((size_t)3 * r->n_angles * sizeof(double));
  if (p.bytes > s->arena_bytes) {
    set_error("mk_server_submit_shadow: output exceeds the budget");
    return MK_ERR_NOMEM;
  }
  p.t0 = Clock::now();
  *id = p.id;
  enqueue(s, std::move(p));
  ++s->st.submitted;
  s->cv_work.notify_one();
  return MK_OK;
}

int mk_server_cancel(mk_server* s, int64_t id, int32_t* h_cancelled) {
  MK_REQUIRE(s && h_cancelled, "mk_server_cancel: null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  MK_REQUIRE(id > 0 && id < s->next_id, "mk_server_cancel: unknown request id %lld", (long long)id);
  *h_cancelled = 0;
  for (auto it = s->waiting.begin(); it != s->waiting.end(); ++it) {
    if (it->id != id) continue;
    s->waiting.erase(it);
    if (s->cancelled_order.size() >= kCancelledKept) {
      s->cancelled.erase(s->cancelled_order.front());
      s->cancelled_order.pop_front();
    }
    s->cancelled.insert(id);
    s->cancelled_order.push_back(id);
    *h_cancelled = 1;
    s->cv_work.notify_one();
    s->cv_done.notify_all();
    break;
  }
  return MK_OK;
}

int mk_server_poll(mk_server* s, int64_t* ids, double* lat, int32_t cap, int32_t* n) {
  MK_REQUIRE(s && n && (cap == 0 || (ids && lat)), "mk_server_poll: null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  const int m = std::min<int>(cap, (int)s->unpolled.size());
  for (int k = 0; k < m; ++k) {
    ids[k] = s->unpolled[k].first;
    lat[k] = s->unpolled[k].second;
  }
  s->unpolled.erase(s->unpolled.begin(), s->unpolled.begin() + m);
  *n = m;
  return server_error(s);
}

int mk_server_wait(mk_server* s, int64_t id, double timeout_s) {
  MK_REQUIRE(s, "mk_server_wait: null server");
  std::unique_lock<std::mutex> lk(s->mu);
  if (id >= s->next_id || id == 0) {
    set_error("mk_server_wait: unknown request id %lld", (long long)id);
    return MK_ERR_ARG;
  }
  // a known id is done once it is neither queued nor in flight, so no
  // per-request record outlives the request
  auto ready = [&] {
    if (s->status != MK_OK) return true;
    if (id < 0) return s->waiting.empty() && s->flight.empty() && s->running_id == 0;
    if (s->running_id == id) return false;
    for (const auto& p : s->waiting)
      if (p.id == id) return false;
    for (const auto& kv : s->flight)
      if (kv.second.id == id) return false;
    return true;
  };
  if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ready)) {
    set_error("mk_server_wait: timed out after %.3f s", timeout_s);
    return MK_ERR_TIMEOUT;
  }
  if (int rc = server_error(s)) return rc;
  if (id > 0 && s->cancelled.count(id)) {
    set_error("mk_server_wait: request %lld was cancelled", (long long)id);
    return MK_ERR_STATE;
  }
  return MK_OK;
}

int mk_server_get_stats(mk_server* s, mk_server_stats* out) {
  MK_REQUIRE(s && out, "mk_server_get_stats: null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  *out = s->st;
  out->queued = (int64_t)s->waiting.size();
  out->in_flight = (int64_t)s->flight.size() + (s->running_id != 0 ? 1 : 0);
  out->hbm_in_use = s->in_use;
  out->lanes = s->L;
  out->steps_per_launch = s->M;
  return server_error(s);
}

int mk_server_destroy(mk_server* s) {
  if (!s) return MK_OK;
  s->stop.store(true);
  s->cv_work.notify_all();
  if (s->worker.joinable()) s->worker.join();
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  if (s->graph) (void)hipGraphExecDestroy(s->graph);
  if (s->ev0) (void)hipEventDestroy(s->ev0);
  if (s->ev1) (void)hipEventDestroy(s->ev1);
  (void)hipFree(s->d_req);
  (void)hipFree(s->d_done);
  (void)hipFree(s->d_segs);
  (void)hipFree(s->d_head);
  (void)hipFree(s->d_f);
  (void)hipFree(s->d_i);
  (void)hipFree(s->d_arena);
  (void)hipFree(s->d_p);
  (void)hipFree(s->d_pdone);
  (void)hipFree(s->d_direct);
  (void)hipFree(s->d_direct_img);
  for (void* h : {(void*)s->p_ctl, (void*)s->p_segs, (void*)s->p_desc, (void*)s->p_fin, (void*)s->p_in,
                  (void*)s->p_out})
    if (h) (void)hipHostFree(h);
  if (s->h_req) (void)hipHostFree(s->h_req);
  if (s->h_segs) (void)hipHostFree(s->h_segs);
  if (s->h_done) (void)hipHostFree(s->h_done);
  if (s->h_tail) (void)hipHostFree(s->h_tail);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
  return MK_OK;
}

}  // extern "C"
