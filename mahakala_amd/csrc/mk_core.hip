// mk_core.hip -- runtime helpers, camera, metric, integrator, shadow
// bisection and transfer kernels of the per-ray hot path (gfx950).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "mk_device.h"
#include "mk_internal.h"

namespace mk {

static thread_local std::string g_error;

void set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_error = buf;
}

constexpr int kBlock = 256;
// Per-ray walks (integrate, select) launch one wave per workgroup: a launch of
// a few thousand rays then spreads over as many CUs as it has waves, and the
// FP64 walk runs faster with fewer waves resident on a CU.
constexpr int kChainBlock = 64;

// ------------------------------------------------------------------ kernels
__global__ void k_camera_grid(double a, double ci, double si, double d, double lo, double step, int res,
                              double* out) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (long)res * res) return;
  const int i = (int)(p / res), j = (int)(p % res);
  double sx[4], sv[4], st[8];
  camera_ray(grid_coord(lo, step, i), grid_coord(lo, step, j), ci, si, d, sx, sv);
  nullify(sx, sv, a, st);
#pragma unroll
  for (int k = 0; k < 8; ++k) out[8 * p + k] = st[k];
}

__global__ void k_camera_points(double a, double ci, double si, double d, const double* px, const double* py, long n,
                                double* out) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  double sx[4], sv[4], st[8];
  camera_ray(px[p], py[p], ci, si, d, sx, sv);
  nullify(sx, sv, a, st);
#pragma unroll
  for (int k = 0; k < 8; ++k) out[8 * p + k] = st[k];
}

__global__ void k_nullify(const double* x, const double* v, long n, double a, double* out) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  double xx[4], vv[4], st[8];
#pragma unroll
  for (int k = 0; k < 4; ++k) { xx[k] = x[4 * p + k]; vv[k] = v[4 * p + k]; }
  nullify(xx, vv, a, st);
#pragma unroll
  for (int k = 0; k < 8; ++k) out[8 * p + k] = st[k];
}

__global__ void k_metric(const double* x, long n, double a, int inverse, double* g) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const KS<double> m = ks_metric<double>(x[4 * p + 1], x[4 * p + 2], x[4 * p + 3], a);
  // covariant l_mu = (1, l_i); contravariant l^mu = (-1, l_i)
  const double l[4] = {inverse ? -1.0 : 1.0, m.l1, m.l2, m.l3};
  const double sgn = inverse ? -1.0 : 1.0;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
      g[16 * p + 4 * i + j] = (i == j ? (i == 0 ? -1.0 : 1.0) : 0.0) + sgn * m.f * (l[i] * l[j]);
}

__global__ void k_rhs(const double* s, long n, double a, double* out) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const double* q = s + 8 * p;
  double acc[4];
  ks_accel<double>(q[1], q[2], q[3], q[4], q[5], q[6], q[7], a, acc);
#pragma unroll
  for (int k = 0; k < 4; ++k) { out[8 * p + k] = q[4 + k]; out[8 * p + 4 + k] = acc[k]; }
}

__global__ void k_rk4(const double* s, const double* dt, long n, double a, double* out) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  double st[8], nw[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) st[k] = s[8 * p + k];
  rk4_step<double>(st, dt[p], a, nw);
#pragma unroll
  for (int k = 0; k < 8; ++k) out[8 * p + k] = nw[k];
}

__global__ void k_radius(const double* x, long n, int stride, double a, double* r) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const double* q = x + (long)stride * p;
  r[p] = radius_cal(q[1], q[2], q[3], a);
}

// Per-ray walk of the reference's fixed-N scan.  Row n holds (s_n, dt_n); a
// ray is frozen from the first row whose step is rejected (dt_n = 0), and
// every later row repeats (s_K, 0).  Also tracks argmax over rows of dt and
// the radius one row before it (select_photons_integrator).
struct WalkOut {
  int K;
  double last[8];
  double sel_r;
};

__device__ __forceinline__ void walk_ray(const double s0[8], int N, double div, double tol, double a, double rEH,
                                         WalkOut& o) {
  double s[8], nw[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) s[k] = s0[k];
  double r = radius_cal(s[1], s[2], s[3], a);
  double dt = step_rule(r, rEH, div, tol);
  double best = -__builtin_inf(), r_prev = __builtin_nan(""), r_at = __builtin_nan("");
  int besti = -1, K = N;
  for (int n = 0; n < N; ++n) {
    bool frozen = (dt == 0.0);
    double rn = 0.0, dtn = 0.0;
    if (!frozen) {
      rk4_step<double>(s, dt, a, nw);
      rn = radius_cal(nw[1], nw[2], nw[3], a);
      dtn = step_rule(rn, rEH, div, tol);
      frozen = (dtn == 0.0);
    }
    const double row_dt = frozen ? 0.0 : dt;
    if (row_dt > best) { best = row_dt; besti = n; r_at = r_prev; }
    if (frozen) { K = n; break; }
    r_prev = r;
    if (n == N - 1) break;
#pragma unroll
    for (int k = 0; k < 8; ++k) s[k] = nw[k];
    dt = dtn;
    r = rn;
  }
  o.K = K;
#pragma unroll
  for (int k = 0; k < 8; ++k) o.last[k] = s[k];
  o.sel_r = (besti >= 1) ? r_at : r;
}

__global__ void __launch_bounds__(256) k_integrate_final(const double* s0, long n, int N, double div, double tol,
                                                         double a, double rEH, int32_t* K, double* last,
                                                         double* sel) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  double st[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) st[k] = s0[8 * p + k];
  WalkOut o;
  walk_ray(st, N, div, tol, a, rEH, o);
  if (K) K[p] = o.K;
  if (last)
#pragma unroll
    for (int k = 0; k < 8; ++k) last[8 * p + k] = o.last[k];
  if (sel) sel[p] = o.sel_r;
}

__global__ void k_max_K(const int32_t* K, long n, int N, int* out /* [0] max K, [1] any K >= N */) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int k = K[p];
  atomicMax(&out[0], k);
  if (k >= N) atomicOr(&out[1], 1);
}

__global__ void __launch_bounds__(256) k_integrate_history(const double* s0, long n, double div, double tol, double a,
                                                           double rEH, long T, double* S, double* D) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  double s[8], nw[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) s[k] = s0[8 * p + k];
  double dt = step_rule(radius_cal(s[1], s[2], s[3], a), rEH, div, tol);
  bool frozen = false;
  for (long row = 0; row < T; ++row) {
    double edt = 0.0, dtn = 0.0;
    if (!frozen) {
      if (dt == 0.0) {
        frozen = true;
      } else {
        rk4_step<double>(s, dt, a, nw);
        dtn = step_rule(radius_cal(nw[1], nw[2], nw[3], a), rEH, div, tol);
        if (dtn == 0.0) frozen = true;
        else edt = dt;
      }
    }
    double* out = S + 8 * (row * n + p);
#pragma unroll
    for (int k = 0; k < 8; ++k) out[k] = s[k];
    D[row * n + p] = edt;
    if (!frozen) {
#pragma unroll
      for (int k = 0; k < 8; ++k) s[k] = nw[k];
      dt = dtn;
    }
  }
}

// select_photons_integrator at (x, y) image points, and the bisection that
// calls it: all angles share one bracket width (geodesics.py:416-433), so
// each lane runs its own angle for the fixed number of halvings.
__global__ void __launch_bounds__(256) k_select(double a, double ci, double si, double d, const double* px,
                                                const double* py, long n, int N, double rEH, double* out) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  double sx[4], sv[4], st[8];
  camera_ray(px[p], py[p], ci, si, d, sx, sv);
  nullify(sx, sv, a, st);
  WalkOut o;
  walk_ray(st, N, 40.0, 1e-2, a, rEH, o);
  out[p] = o.sel_r;
}

__global__ void __launch_bounds__(256) k_bisect(double a, double ci, double si, double d, const double* cang,
                                                const double* sang, long n, int N, int iters, double rEH,
                                                double* radii) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  double inner = 0.5, outer = 10.0;
  for (int it = 0; it < iters; ++it) {
    const double mid = (outer - inner) / 2 + inner;
    double sx[4], sv[4], st[8];
    camera_ray(cang[p] * mid, sang[p] * mid, ci, si, d, sx, sv);
    nullify(sx, sv, a, st);
    WalkOut o;
    walk_ray(st, N, 40.0, 1e-2, a, rEH, o);
    if (o.sel_r < 100.0) inner = mid;
    else outer = mid;
  }
  radii[p] = inner;
}

// The same bisection, K halvings per round: each angle's group of 2^K lanes
// integrates all 2^K - 1 midpoints the next K halvings could visit (the
// nodes of a binary tree in heap order, each computed along its path with
// the reference's (outer - inner) / 2 + inner), then one lane replays the K
// halvings from the classifications.  The halvings, midpoints and result are
// the sequential bisection's, in ceil(iters / K) integration rounds instead
// of iters.
__global__ void __launch_bounds__(1024) k_bisect_tree(double a, double ci, double si, double d, const double* cang,
                                                      const double* sang, long n, int N, int iters, double rEH,
                                                      int K, double* radii) {
  __shared__ double br[2 * 1024];  // inner, outer per angle of the block
  __shared__ int cls[1024];
  const int gsz = 1 << K, G = blockDim.x >> K;
  const int g = threadIdx.x >> K, j = threadIdx.x & (gsz - 1);
  const long p = (long)blockIdx.x * G + g;
  if (j == 0) {
    br[2 * g] = 0.5;
    br[2 * g + 1] = 10.0;
  }
  __syncthreads();
  for (int done = 0; done < iters;) {
    const int kr = min(K, iters - done);
    int c = 0;
    if (p < n && j < (1 << kr) - 1) {
      // node j (heap order): the bits of j + 1 below its leading one are the
      // path from the root, 1 = the ray fell in (inner = mid)
      const int m = j + 1, depth = 31 - __builtin_clz(m);
      double lo = br[2 * g], hi = br[2 * g + 1];
      double mid = (hi - lo) / 2 + lo;
      for (int b = depth - 1; b >= 0; --b) {
        if ((m >> b) & 1) lo = mid;
        else hi = mid;
        mid = (hi - lo) / 2 + lo;
      }
      double sx[4], sv[4], st[8];
      camera_ray(cang[p] * mid, sang[p] * mid, ci, si, d, sx, sv);
      nullify(sx, sv, a, st);
      WalkOut o;
      walk_ray(st, N, 40.0, 1e-2, a, rEH, o);
      c = o.sel_r < 100.0 ? 1 : 0;
    }
    cls[threadIdx.x] = c;
    __syncthreads();
    if (j == 0 && p < n) {
      double lo = br[2 * g], hi = br[2 * g + 1];
      int node = 0;
      for (int l = 0; l < kr; ++l) {
        const double mid = (hi - lo) / 2 + lo;
        if (cls[(g << K) + node]) {
          lo = mid;
          node = 2 * node + 2;
        } else {
          hi = mid;
          node = 2 * node + 1;
        }
      }
      br[2 * g] = lo;
      br[2 * g + 1] = hi;
    }
    __syncthreads();
    done += kr;
  }
  if (j == 0 && p < n) radii[p] = br[2 * g];
}

// Bracket indices [b0, b1] at level l (after l halvings the brackets are
// 9.5 / 2^l wide from 0.5, index bits = the halvings' outcomes, 1 = fell)
// that meet the band [rho - e, rho + e]; empty without a prediction.  Used
// only to choose which midpoints to integrate: a wrong choice costs time,
// never a result.
__device__ __forceinline__ long long band_range(double rho, double e, int l, long long& b0) {
  if (!(rho > 0.0 && rho < 100.0)) {  // NaN, or nothing a bracket in [0.5, 10] could meet
    b0 = 0;
    return 0;
  }
  const long long nb = 1ll << l;
  const double w = 9.5 / (double)nb;
  b0 = (long long)__builtin_floor((rho - e - 0.5) / w);
  long long b1 = (long long)__builtin_floor((rho + e - 0.5) / w);
  if (b0 < 0) b0 = 0;
  if (b1 > nb - 1) b1 = nb - 1;
  return b1 >= b0 ? b1 - b0 + 1 : 0;
}
constexpr int kBandMaxLevels = 48;

// Midpoints in the band over levels [0, L).
__device__ __forceinline__ long long band_nodes(double rho, double e, int L) {
  long long t = 0, b0;
  for (int l = 0; l < L; ++l) t += band_range(rho, e, l, b0);
  return t;
}

// The bisection in fewer rounds, guided by a predicted radius per angle
// (radii[p] on entry, NaN for none): each angle's B lanes first integrate, in
// one round, every midpoint whose bracket meets a band about the prediction:
// the widest band of half-width up to e whose midpoints of every level fit
// in B lanes with one angle per block (several angles per block: the band
// of half-width e over as many levels as fit).
// One lane then replays the halvings from the classifications while the path
// stays in the band; the halvings left (path out of the band, or levels
// beyond it) run in tree rounds as in k_bisect_tree.  Every midpoint is
// computed along its path with the reference's (outer - inner) / 2 + inner
// and the replay uses only classifications of the midpoints the sequential
// bisection visits, so the result is its result, bit for bit, whatever the
// prediction; the prediction decides only how many rounds it takes.  G
// angles per block.
template <int B, int G>
__global__ void __launch_bounds__(B * G) k_bisect_band(double a, double ci, double si, double d, const double* cang,
                                                       const double* sang, long n, int N, int iters, double rEH,
                                                       double e, double* radii) {
  constexpr int KB = B >= 255 ? 8 : B >= 127 ? 7 : B >= 63 ? 6 : B >= 31 ? 5 : 4;  // 2^KB - 1 <= B
  __shared__ double br[2 * G];
  __shared__ int cls[B * G], resolved[G];
  const int g = threadIdx.x / B, j = threadIdx.x % B;
  const long p = (long)blockIdx.x * G + g;
  const bool valid = p < n;
  const double rho = valid ? radii[p] : __builtin_nan("");
  const int lmax = iters < kBandMaxLevels ? iters : kBandMaxLevels;
  // one angle per block: narrow the band until every level fits (several
  // angles per block keep the band and fewer levels: their lanes are few)
  if (G == 1 && valid && band_nodes(rho, e, lmax) > B) {
    double lo = 0.0, hi = e;
    for (int it = 0; it < 30; ++it) {
      const double m = 0.5 * (lo + hi);
      if (band_nodes(rho, m, lmax) <= B) lo = m;
      else hi = m;
    }
    e = lo;
  }
  // this lane's midpoint in the band: level ml, bracket mb
  int L = 0, ml = -1;
  long long mb = 0;
  if (valid) {
    long long total = 0;
    for (int l = 0; l < lmax; ++l) {
      long long b0;
      const long long cnt = band_range(rho, e, l, b0);
      if (cnt == 0 || total + cnt > B) break;
      if (j >= total && j < total + cnt) {
        ml = l;
        mb = b0 + (j - total);
      }
      total += cnt;
      L = l + 1;
    }
  }
  int c = 0;
  if (ml >= 0) {
    double lo = 0.5, hi = 10.0, mid = (hi - lo) / 2 + lo;
    for (int k = ml - 1; k >= 0; --k) {
      if ((mb >> k) & 1) lo = mid;
      else hi = mid;
      mid = (hi - lo) / 2 + lo;
    }
    double sx[4], sv[4], st[8];
    camera_ray(cang[p] * mid, sang[p] * mid, ci, si, d, sx, sv);
    nullify(sx, sv, a, st);
    WalkOut o;
    walk_ray(st, N, 40.0, 1e-2, a, rEH, o);
    c = o.sel_r < 100.0 ? 1 : 0;
  }
  cls[threadIdx.x] = c;
  __syncthreads();
  if (j == 0) {
    double lo = 0.5, hi = 10.0;
    long long b = 0, off = 0;
    int l = 0;
    for (; l < L; ++l) {
      long long b0;
      const long long cnt = band_range(rho, e, l, b0);
      if (b < b0 || b >= b0 + cnt) break;  // the path left the band
      const double mid = (hi - lo) / 2 + lo;
      if (cls[g * B + (int)(off + b - b0)]) {
        lo = mid;
        b = 2 * b + 1;
      } else {
        hi = mid;
        b = 2 * b;
      }
      off += cnt;
    }
    br[2 * g] = lo;
    br[2 * g + 1] = hi;
    resolved[g] = valid ? l : iters;
  }
  __syncthreads();
  // the halvings left, KB or fewer per round (spread evenly over the rounds);
  // every lane reads its angle's state after the barrier that follows its
  // update
  while (true) {
    const int rem = iters - resolved[g];
    if (!__syncthreads_or(rem > 0)) break;
    const int rounds = (rem + KB - 1) / KB;
    const int kr = rem > 0 ? (rem + rounds - 1) / rounds : 0;
    c = 0;
    if (j < (1 << kr) - 1) {
      const int m = j + 1, depth = 31 - __builtin_clz(m);
      double lo = br[2 * g], hi = br[2 * g + 1];
      double mid = (hi - lo) / 2 + lo;
      for (int k = depth - 1; k >= 0; --k) {
        if ((m >> k) & 1) lo = mid;
        else hi = mid;
        mid = (hi - lo) / 2 + lo;
      }
      double sx[4], sv[4], st[8];
      camera_ray(cang[p] * mid, sang[p] * mid, ci, si, d, sx, sv);
      nullify(sx, sv, a, st);
      WalkOut o;
      walk_ray(st, N, 40.0, 1e-2, a, rEH, o);
      c = o.sel_r < 100.0 ? 1 : 0;
    }
    cls[threadIdx.x] = c;
    __syncthreads();
    if (j == 0 && kr > 0) {
      double lo = br[2 * g], hi = br[2 * g + 1];
      int node = 0;
      for (int l = 0; l < kr; ++l) {
        const double mid = (hi - lo) / 2 + lo;
        if (cls[g * B + node]) {
          lo = mid;
          node = 2 * node + 2;
        } else {
          hi = mid;
          node = 2 * node + 1;
        }
      }
      br[2 * g] = lo;
      br[2 * g + 1] = hi;
      resolved[g] += kr;
    }
    __syncthreads();
  }
  if (j == 0 && valid) radii[p] = br[2 * g];
}

__global__ void k_synchrotron(const double* Ne, const double* Th, const double* B, const double* pitch,
                              const double* nu, long n, int invariant, double rescale, double* j, double* al) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  double jj, aa;
  synchrotron<double>(Ne[p], Th[p], B[p], pitch[p], nu[p], invariant != 0, rescale, jj, aa);
  j[p] = jj;
  al[p] = aa;
}

// One thread per ray walks its rows in the reference's order.  The loads of 8
// rows are issued before their (sequential) arithmetic, so a thread keeps 24
// loads in flight instead of 3: the same arithmetic, bit for bit.
constexpr int kSolveRows = 8;

__global__ void k_solve(const double* j, const double* al, const double* dt, long T, long npx, double L, double* I,
                        double* dI) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npx) return;
  double Iv = 0.0;
  long i = T - 1;
  for (; i >= kSolveRows; i -= kSolveRows) {
    double jj[kSolveRows], aa[kSolveRows], dd[kSolveRows];
#pragma unroll
    for (int u = 0; u < kSolveRows; ++u) {
      jj[u] = j[(i - u) * npx + p];
      aa[u] = al[(i - u) * npx + p];
      dd[u] = dt[(i - u - 1) * npx + p];
    }
#pragma unroll
    for (int u = 0; u < kSolveRows; ++u) {
      const double d = -dd[u] * L * (jj[u] - (aa[u] * Iv));
      Iv = Iv + d;
      if (dI) dI[(T - 1 - (i - u)) * npx + p] = d;
    }
  }
  for (; i >= 1; --i) {
    const double d = -dt[(i - 1) * npx + p] * L * (j[i * npx + p] - (al[i * npx + p] * Iv));
    Iv = Iv + d;
    if (dI) dI[(T - 1 - i) * npx + p] = d;
  }
  I[p] = Iv;
}

__global__ void k_attenuated(const double* j, const double* al, const double* dt, long T, long npx, double L,
                             double* out) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npx) return;
  double tau = 0.0;
  long i = 1;
  for (; i + kSolveRows <= T; i += kSolveRows) {  // loads of 8 rows first, as k_solve
    double jj[kSolveRows], aa[kSolveRows], dd[kSolveRows];
#pragma unroll
    for (int u = 0; u < kSolveRows; ++u) {
      jj[u] = j[(i + u) * npx + p];
      aa[u] = al[(i + u) * npx + p];
      dd[u] = dt[(i + u - 1) * npx + p];
    }
#pragma unroll
    for (int u = 0; u < kSolveRows; ++u) {
      const double src = -jj[u] * dd[u] * L;
      const double dtau = aa[u] * dd[u] * L;
      out[(i + u - 1) * npx + p] = ::exp(-tau) * src;
      tau = tau - dtau;
    }
  }
  for (; i < T; ++i) {
    const double src = -j[i * npx + p] * dt[(i - 1) * npx + p] * L;
    const double dtau = al[i * npx + p] * dt[(i - 1) * npx + p] * L;
    out[(i - 1) * npx + p] = ::exp(-tau) * src;
    tau = tau - dtau;
  }
}

__global__ void k_rlow_rhigh(const double* dens, const double* u, const double* beta, long n, double rl, double rh,
                             double ge, double gi, double* out) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  out[p] = rlow_rhigh<double>(dens[p], u[p], beta[p], rl, rh, ge, gi);
}

__global__ void k_copy16(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

double horizon(double a) { return 1.0 + std::sqrt(1.0 - a * a); }

}  // namespace mk

using namespace mk;

// ------------------------------------------------ allocation-free staging
// mk_shadow_bisect_noalloc keeps its staging (cos and sin of the angles in,
// radii out) in this fixed array instead of allocating: each GPU holds its
// own copy (1.5 MiB) until the process ends.  Host code reaches it only
// through the per-device address from hipGetSymbolAddress; the array's name
// in host code is not a device address.  The host side of the staging is a
// fixed host array too.  One lock serialises the calls, so calls from
// several host threads do not share the array at once.
__device__ __align__(256) double scratch[3 * MK_NOALLOC_MAX_ANGLES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
double* g_noalloc_dev[kMaxDevices] = {};
double g_noalloc_host[2 * MK_NOALLOC_MAX_ANGLES];

// The current device's copy of `scratch`, looked up once per device (called
// with g_noalloc_mu held).
int noalloc_scratch(double** out) {
  int dev = 0;
  MK_HIP(hipGetDevice(&dev));
  MK_REQUIRE(dev >= 0 && dev < kMaxDevices, "mk_shadow_bisect_noalloc: device %d beyond %d", dev, kMaxDevices);
  if (!g_noalloc_dev[dev]) {
    void* ptr = nullptr;
    MK_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_noalloc_dev[dev] = static_cast<double*>(ptr);
  }
  *out = g_noalloc_dev[dev];
  return MK_OK;
}
}  // namespace

// =================================================================== C ABI
extern "C" {

const char* mk_last_error(void) { return g_error.c_str(); }
int mk_version(void) { return 10000; }

int mk_device_info(char* name, size_t len, int* cus, size_t* total) {
  int dev = 0;
  MK_HIP(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  MK_HIP(hipGetDeviceProperties(&prop, dev));
  if (name && len) {
    std::strncpy(name, prop.gcnArchName, len - 1);
    name[len - 1] = 0;
  }
  if (cus) *cus = prop.multiProcessorCount;
  if (total) *total = prop.totalGlobalMem;
  return MK_OK;
}

int mk_device_mem(size_t* free_bytes, size_t* total_bytes) {
  MK_REQUIRE(free_bytes && total_bytes, "mk_device_mem: null argument");
  MK_HIP(hipMemGetInfo(free_bytes, total_bytes));
  return MK_OK;
}

int mk_malloc(void** p, size_t bytes) {
  MK_REQUIRE(p != nullptr, "mk_malloc: null out pointer");
  *p = nullptr;
  if (bytes == 0) return MK_OK;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    set_error("mk_malloc(%zu): %s", bytes, hipGetErrorString(e));
    return MK_ERR_NOMEM;
  }
  return MK_OK;
}
int mk_free(void* p) {
  if (p) MK_HIP(hipFree(p));
  return MK_OK;
}
int mk_memcpy_h2d(void* d, const void* h, size_t bytes, void* s) {
  if (!bytes) return MK_OK;
  MK_HIP(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, as_stream(s)));
  MK_HIP(hipStreamSynchronize(as_stream(s)));
  return MK_OK;
}
int mk_memcpy_d2h(void* h, const void* d, size_t bytes, void* s) {
  if (!bytes) return MK_OK;
  MK_HIP(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, as_stream(s)));
  MK_HIP(hipStreamSynchronize(as_stream(s)));
  return MK_OK;
}
int mk_memset(void* d, int v, size_t bytes, void* s) {
  if (!bytes) return MK_OK;
  MK_HIP(hipMemsetAsync(d, v, bytes, as_stream(s)));
  return MK_OK;
}
int mk_device_copy(void* d, const void* src, size_t bytes, void* s) {
  MK_REQUIRE(d && src && bytes % 16 == 0, "mk_device_copy: pointers must be non-null and bytes a multiple of 16");
  if (!bytes) return MK_OK;
  const size_t n = bytes / 16;
  const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 256 * 64);
  k_copy16<<<blocks, 256, 0, as_stream(s)>>>((const uint4*)src, (uint4*)d, n);
  MK_LAUNCH_CHECK();
  return MK_OK;
}
int mk_stream_create(void** s) {
  hipStream_t st;
  MK_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  *s = st;
  return MK_OK;
}
int mk_stream_destroy(void* s) {
  if (s) MK_HIP(hipStreamDestroy(as_stream(s)));
  return MK_OK;
}
int mk_stream_sync(void* s) {
  MK_HIP(hipStreamSynchronize(as_stream(s)));
  return MK_OK;
}
int mk_event_create(void** e) {
  hipEvent_t ev;
  MK_HIP(hipEventCreate(&ev));
  *e = ev;
  return MK_OK;
}
int mk_event_destroy(void* e) {
  if (e) MK_HIP(hipEventDestroy((hipEvent_t)e));
  return MK_OK;
}
int mk_event_record(void* e, void* s) {
  MK_HIP(hipEventRecord((hipEvent_t)e, as_stream(s)));
  return MK_OK;
}
int mk_event_sync(void* e) {
  MK_HIP(hipEventSynchronize((hipEvent_t)e));
  return MK_OK;
}
int mk_event_elapsed_ms(void* e0, void* e1, float* ms) {
  MK_HIP(hipEventElapsedTime(ms, (hipEvent_t)e0, (hipEvent_t)e1));
  return MK_OK;
}

static void cam_angles(double inc, double* ci, double* si) {
  const double i = inc * M_PI / 180;
  *ci = std::cos(i);
  *si = std::sin(i);
}

int mk_camera_grid(double a, double inc, double dist, double lo, double hi, int res, double* d_out, void* s) {
  MK_REQUIRE(res > 0 && d_out, "mk_camera_grid: bad arguments");
  double ci, si;
  cam_angles(inc, &ci, &si);
  const double step = (hi - lo) / (double)(2 * res);
  const long n = (long)res * res;
  k_camera_grid<<<grid_for(n, kBlock), kBlock, 0, as_stream(s)>>>(a, ci, si, dist, lo, step, res, d_out);
  MK_LAUNCH_CHECK();
  return MK_OK;
}

// host (x, y) = (cos(angle) radius, sin(angle) radius), as get_camera_pixel
static int upload_xy(const double* radius, const double* angle, long n, double** dx, double** dy, hipStream_t st) {
  std::vector<double> hx(n), hy(n);
  for (long p = 0; p < n; ++p) {
    hx[p] = std::cos(angle[p]) * radius[p];
    hy[p] = std::sin(angle[p]) * radius[p];
  }
  MK_HIP(hipMalloc(dx, n * sizeof(double)));
  MK_HIP(hipMalloc(dy, n * sizeof(double)));
  MK_HIP(hipMemcpyAsync(*dx, hx.data(), n * sizeof(double), hipMemcpyHostToDevice, st));
  MK_HIP(hipMemcpyAsync(*dy, hy.data(), n * sizeof(double), hipMemcpyHostToDevice, st));
  MK_HIP(hipStreamSynchronize(st));
  return MK_OK;
}

int mk_camera_pixels(double a, double inc, double dist, const double* radius, const double* angle, long n,
                     double* d_out, void* s) {
  MK_REQUIRE(n >= 0 && radius && angle && d_out, "mk_camera_pixels: bad arguments");
  if (n == 0) return MK_OK;
  double ci, si;
  cam_angles(inc, &ci, &si);
  double *dx = nullptr, *dy = nullptr;
  int rc = upload_xy(radius, angle, n, &dx, &dy, as_stream(s));
  if (rc == MK_OK) {
    k_camera_points<<<grid_for(n, kBlock), kBlock, 0, as_stream(s)>>>(a, ci, si, dist, dx, dy, n, d_out);
    hipError_t e = hipStreamSynchronize(as_stream(s));
    if (e != hipSuccess) { set_error("mk_camera_pixels: %s", hipGetErrorString(e)); rc = MK_ERR_HIP; }
  }
  (void)hipFree(dx);
  (void)hipFree(dy);
  return rc;
}

int mk_initial_condition(const double* d_x, const double* d_v, long n, double a, double* d_out, void* s) {
  MK_REQUIRE(n >= 0 && d_x && d_v && d_out, "mk_initial_condition: bad arguments");
  if (!n) return MK_OK;
  k_nullify<<<grid_for(n, kBlock), kBlock, 0, as_stream(s)>>>(d_x, d_v, n, a, d_out);
  MK_LAUNCH_CHECK();
  return MK_OK;
}

int mk_metric(const double* d_x, long n, double a, double* d_g, void* s) {
  MK_REQUIRE(n >= 0 && d_x && d_g, "mk_metric: bad arguments");
  if (!n) return MK_OK;
  k_metric<<<grid_for(n, kBlock), kBlock, 0, as_stream(s)>>>(d_x, n, a, 0, d_g);
  MK_LAUNCH_CHECK();
  return MK_OK;
}
int mk_imetric(const double* d_x, long n, double a, double* d_g, void* s) {
  MK_REQUIRE(n >= 0 && d_x && d_g, "mk_imetric: bad arguments");
  if (!n) return MK_OK;
  k_metric<<<grid_for(n, kBlock), kBlock, 0, as_stream(s)>>>(d_x, n, a, 1, d_g);
  MK_LAUNCH_CHECK();
  return MK_OK;
}
int mk_rhs(const double* d_s, long n, double a, double* d_out, void* s) {
  MK_REQUIRE(n >= 0 && d_s && d_out, "mk_rhs: bad arguments");
  if (!n) return MK_OK;
  k_rhs<<<grid_for(n, kBlock), kBlock, 0, as_stream(s)>>>(d_s, n, a, d_out);
  MK_LAUNCH_CHECK();
  return MK_OK;
}
int mk_rk4(const double* d_s, const double* d_dt, long n, double a, double* d_out, void* s) {
  MK_REQUIRE(n >= 0 && d_s && d_dt && d_out, "mk_rk4: bad arguments");
  if (!n) return MK_OK;
  k_rk4<<<grid_for(n, kBlock), kBlock, 0, as_stream(s)>>>(d_s, d_dt, n, a, d_out);
  MK_LAUNCH_CHECK();
  return MK_OK;
}
int mk_radius_cal(const double* d_x, long n, int stride, double a, double* d_r, void* s) {
  MK_REQUIRE(n >= 0 && stride >= 4 && d_x && d_r, "mk_radius_cal: bad arguments");
  if (!n) return MK_OK;
  k_radius<<<grid_for(n, kBlock), kBlock, 0, as_stream(s)>>>(d_x, n, stride, a, d_r);
  MK_LAUNCH_CHECK();
  return MK_OK;
}

int mk_integrate_final(const double* d_s0, long n, int N, double div, double tol, double a, int32_t* d_K,
                       double* d_last, double* d_sel, void* s) {
  MK_REQUIRE(n >= 0 && N >= 1 && d_s0 && div != 0.0, "mk_integrate_final: bad arguments");
  if (!n) return MK_OK;
  k_integrate_final<<<grid_for(n, kChainBlock), kChainBlock, 0, as_stream(s)>>>(d_s0, n, N, div, tol, a, horizon(a), d_K,
                                                                      d_last, d_sel);
  MK_LAUNCH_CHECK();
  return MK_OK;
}

int mk_trim_rows(const int32_t* d_K, long n, int N, long* h_T, void* s) {
  MK_REQUIRE(n >= 0 && N >= 1 && h_T, "mk_trim_rows: bad arguments");
  if (n == 0) {
    *h_T = N;
    return MK_OK;
  }
  int* d = nullptr;
  MK_HIP(hipMalloc(&d, 2 * sizeof(int)));
  int h[2] = {0, 0};
  hipError_t e = hipMemcpyAsync(d, h, sizeof h, hipMemcpyHostToDevice, as_stream(s));
  if (e == hipSuccess) {
    k_max_K<<<grid_for(n, kBlock), kBlock, 0, as_stream(s)>>>(d_K, n, N, d);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(h, d, sizeof h, hipMemcpyDeviceToHost, as_stream(s));
  if (e == hipSuccess) e = hipStreamSynchronize(as_stream(s));
  (void)hipFree(d);
  if (e != hipSuccess) {
    set_error("mk_trim_rows: %s", hipGetErrorString(e));
    return MK_ERR_HIP;
  }
  // argmax of the all-frozen row mask, "+2" slice, "< 1 -> N" rule
  const long first = h[1] ? 0 : h[0];
  *h_T = (first < 1) ? N : (first + 2 < N ? first + 2 : N);
  return MK_OK;
}

int mk_integrate_history(const double* d_s0, long n, int N, double div, double tol, double a, long T, double* d_S,
                         double* d_dt, void* s) {
  MK_REQUIRE(n >= 0 && T >= 0 && T <= N && d_s0 && d_S && d_dt, "mk_integrate_history: bad arguments");
  if (!n || !T) return MK_OK;
  k_integrate_history<<<grid_for(n, kChainBlock), kChainBlock, 0, as_stream(s)>>>(d_s0, n, div, tol, a, horizon(a), T, d_S,
                                                                        d_dt);
  MK_LAUNCH_CHECK();
  return MK_OK;
}

int mk_select_photons(double inc, const double* h_angle, const double* h_radius, long n, double a, double dist,
                      int max_steps, double* h_out) {
  MK_REQUIRE(n >= 0 && max_steps >= 1 && h_angle && h_radius && h_out, "mk_select_photons: bad arguments");
  if (!n) return MK_OK;
  double ci, si;
  cam_angles(inc, &ci, &si);
  double *dx = nullptr, *dy = nullptr, *dout = nullptr;
  int rc = upload_xy(h_radius, h_angle, n, &dx, &dy, nullptr);
  hipError_t e = hipSuccess;
  if (rc == MK_OK) e = hipMalloc(&dout, n * sizeof(double));
  if (rc == MK_OK && e == hipSuccess) {
    k_select<<<grid_for(n, kChainBlock), kChainBlock>>>(a, ci, si, dist, dx, dy, n, max_steps, horizon(a), dout);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpy(h_out, dout, n * sizeof(double), hipMemcpyDeviceToHost);
  }
  (void)hipFree(dx);
  (void)hipFree(dy);
  (void)hipFree(dout);
  if (rc == MK_OK && e != hipSuccess) {
    set_error("mk_select_photons: %s", hipGetErrorString(e));
    rc = MK_ERR_HIP;
  }
  return rc;
}

// Halvings of a bisection with the reference's shared bracket
// (geodesics.py:417-431): every angle does the same number.
static int bisect_iterations(double error_allowed, int max_it) {
  int iters = 0;
  double inner = 0.5, outer = 10.0;
  while (outer - inner > error_allowed && iters < max_it) {
    inner = (outer - inner) / 2 + inner;
    ++iters;
  }
  return iters;
}

// Half-width (M) of the band about the predicted shadow radius that the
// bisection integrates in its first round: MK_BISECT_BAND in the environment,
// default 0.05 M (the predictions below are within 0.005 M of the bisection's
// result at moderate spins and 0.03 M at a = 0.99, i = 85 deg); 0 turns the
// band off.
static double bisect_band() {
  const char* v = std::getenv("MK_BISECT_BAND");
  if (!v || !*v) return 0.05;
  const double e = std::atof(v);
  return e > 0.0 && e < 10.0 ? e : 0.0;
}

// Critical curve of a distant camera (Bardeen 1973; image x is alpha, image y
// is beta in this camera's convention) as a polar table about the image
// origin: kPolar bins of the distance from the origin to the curve along the
// bin centre's direction, NaN where the curve does not cross.
constexpr int kPolar = 2048;
static void critical_curve_about_origin(double a, double inc_deg, std::vector<double>& R) {
  constexpr int kPts = 1024, kScan = 1024;
  static const std::vector<double> dir = [] {  // cos, sin of the bin centres
    std::vector<double> d(2 * kPolar);
    for (int k = 0; k < kPolar; ++k) {
      const double f = -M_PI + (k + 0.5) * 2.0 * M_PI / kPolar;
      d[2 * k] = std::cos(f);
      d[2 * k + 1] = std::sin(f);
    }
    return d;
  }();
  const double th = std::min(std::max(inc_deg, 1.0), 179.0) * M_PI / 180.0;
  std::vector<double> px, py;
  if (std::fabs(a) < 1e-3) {
    for (int q = 0; q < 2 * kPts; ++q) {
      px.push_back(std::sqrt(27.0) * std::cos(M_PI * q / kPts));
      py.push_back(std::sqrt(27.0) * std::sin(M_PI * q / kPts));
    }
  } else {
    const double aa = std::max(-0.9999, std::min(0.9999, a));
    const double r1 = 2.0 * (1.0 + std::cos(2.0 / 3.0 * std::acos(-std::fabs(aa))));
    const double r2 = 2.0 * (1.0 + std::cos(2.0 / 3.0 * std::acos(std::fabs(aa))));
    const double ct = std::cos(th) / std::sin(th), c2 = std::cos(th) * std::cos(th);
    auto xi_of = [&](double r) { return (r * r * (3.0 - r) - aa * aa * (r + 1.0)) / (aa * (r - 1.0)); };
    auto beta2 = [&](double r) {
      const double xi = xi_of(r);
      const double eta = r * r * r * (4.0 * aa * aa - r * (r - 3.0) * (r - 3.0)) / (aa * aa * (r - 1.0) * (r - 1.0));
      return eta + aa * aa * c2 - xi * xi * ct * ct;
    };
    // the orbit radii the camera sees (beta^2 >= 0: one interval), found on a
    // scan and refined by bisection at both ends
    const double rmin = std::min(r1, r2), rmax = std::max(r1, r2);
    int q0 = -1, q1 = -1;
    for (int q = 0; q <= kScan; ++q) {
      const double r = rmin + (rmax - rmin) * q / kScan;
      if (std::fabs(r - 1.0) > 1e-9 && beta2(r) >= 0.0) {
        if (q0 < 0) q0 = q;
        q1 = q;
      }
    }
    if (q0 >= 0 && q1 > q0) {
      auto edge = [&](int in, int out) {  // the sign change of beta^2 between two scan points
        double u = rmin + (rmax - rmin) * in / kScan, v = rmin + (rmax - rmin) * out / kScan;
        for (int it = 0; it < 60; ++it) {
          const double w = 0.5 * (u + v);
          if (beta2(w) >= 0.0) u = w;
          else v = w;
        }
        return u;
      };
      const double lo = q0 > 0 ? edge(q0, q0 - 1) : rmin, hi = q1 < kScan ? edge(q1, q1 + 1) : rmax;
      // beta ~ sqrt(r - lo) at the ends: samples clustered there (r - lo ~ u^2)
      // are spread evenly along the curve; upper branch, then the lower one back
      std::vector<double> al, be;
      for (int q = 0; q <= kPts; ++q) {
        const double r = lo + (hi - lo) * 0.5 * (1.0 - std::cos(M_PI * q / kPts));
        if (std::fabs(r - 1.0) < 1e-9) continue;
        al.push_back(-xi_of(r) / std::sin(th));
        be.push_back(std::sqrt(std::max(beta2(r), 0.0)));
      }
      for (size_t q = 0; q < al.size(); ++q) {
        px.push_back(al[q]);
        py.push_back(be[q]);
      }
      for (size_t q = al.size(); q-- > 0;) {
        px.push_back(al[q]);
        py.push_back(-be[q]);
      }
    }
  }
  R.assign(kPolar, std::nan(""));
  const size_t m = px.size();
  if (m < 8) return;
  std::vector<double> ph(m);
  for (size_t q = 0; q < m; ++q) ph[q] = std::atan2(py[q], px[q]);
  // each edge's intersections with the rays at the bin centres it spans
  for (size_t q = 0; q < m; ++q) {
    const size_t u = (q + 1) % m;
    const double ax = px[q], ay = py[q], ex = px[u] - ax, ey = py[u] - ay;
    double df = ph[u] - ph[q];
    if (df > M_PI) df -= 2.0 * M_PI;
    if (df < -M_PI) df += 2.0 * M_PI;
    const double f0 = std::min(ph[q], ph[q] + df), f1 = std::max(ph[q], ph[q] + df);
    for (long k = (long)std::ceil((f0 + M_PI) / (2.0 * M_PI) * kPolar - 0.5);
         k <= (long)std::floor((f1 + M_PI) / (2.0 * M_PI) * kPolar - 0.5); ++k) {
      const int kb = (int)(((k % kPolar) + kPolar) % kPolar);
      const double dx = dir[2 * kb], dy = dir[2 * kb + 1], den = dx * ey - dy * ex;
      if (den == 0.0) continue;
      const double t = (ax * ey - ay * ex) / den;
      if (t > 0.0 && !(R[kb] >= t)) R[kb] = t;
    }
  }
}

// Predicted shadow radius along each image direction (angle ang[p], the
// image x axis at 0): the critical curve's table, interpolated between bin
// centres; NaN where the table has no value.  Only the bisection's schedule depends on it.  The
// tables of the last few (spin, inclination) pairs are kept.
namespace {
std::mutex g_curve_mu;
std::vector<std::pair<std::pair<double, double>, std::vector<double>>> g_curves;  // most recent last
constexpr size_t kCurvesKept = 8;
}  // namespace

static void predicted_shadow_radii(double a, double inc_deg, const double* ang, long n, double* rho) {
  std::vector<double> R;
  {
    std::lock_guard<std::mutex> lk(g_curve_mu);
    for (const auto& e : g_curves)
      if (e.first.first == a && e.first.second == inc_deg) R = e.second;
  }
  if (R.empty()) {
    critical_curve_about_origin(a, inc_deg, R);
    std::lock_guard<std::mutex> lk(g_curve_mu);
    if (g_curves.size() >= kCurvesKept) g_curves.erase(g_curves.begin());
    g_curves.push_back({{a, inc_deg}, R});
  }
  for (long p = 0; p < n; ++p) {
    const double f = ang[p] - 2.0 * M_PI * std::floor((ang[p] + M_PI) / (2.0 * M_PI));  // in [-pi, pi)
    const double u = (f + M_PI) / (2.0 * M_PI) * kPolar - 0.5;
    const double fl = std::floor(u);
    const int k0 = (int)((((long)fl % kPolar) + kPolar) % kPolar), k1 = (k0 + 1) % kPolar;
    const double w = u - fl;
    rho[p] = (1.0 - w) * R[k0] + w * R[k1];  // NaN when either bin is empty
  }
}

// Launches the bisection of n angles (cos and sin in dc, ds; device) into dr
// on `st`.  h_rho (n host doubles, may be null): the predicted radii, staged
// through dr for the band form (copied before the call returns).
static hipError_t launch_bisect(double a, double inc, const double* dc, const double* ds, long n, int max_steps,
                                int iters, double dist, double* dr, const double* h_rho, hipStream_t st) {
  double ci, si;
  cam_angles(inc, &ci, &si);
  // The band form.  The FP64 walk slows with the
  // waves resident on a CU, so up to 1024 angles each angle is a block of
  // its own: one wave (the band then fits about 0.012 M about the
  // prediction, which is within 0.007 M of the result for |a| <= 0.9) or, for
  // higher spins, where the prediction drifts further (0.02 M at 0.97, more
  // at high inclination near a = 1), two waves (about 0.03 M).  From 1025
  // to 4096 angles, 8 angles of 32 lanes per block (2048 angles: 6.9 against
  // 9.5 ms in tree rounds, 4096 angles 10.1 against 12.8 ms; 16 lanes per
  // angle: 6.9 and 10.6 ms).
  const double e = bisect_band();
  if (h_rho && e > 0.0 && iters > 1 && n <= 4096) {
    hipError_t err = hipMemcpyAsync(dr, h_rho, n * sizeof(double), hipMemcpyHostToDevice, st);
    if (err == hipSuccess) err = hipStreamSynchronize(st);  // h_rho may be freed after the return
    if (err != hipSuccess) return err;
    const double rEH = horizon(a);
    if (n <= 1024 && std::fabs(a) <= 0.9)
      k_bisect_band<64, 1><<<(unsigned)n, 64, 0, st>>>(a, ci, si, dist, dc, ds, n, max_steps, iters, rEH, e, dr);
    else if (n <= 1024)
      k_bisect_band<128, 1><<<(unsigned)n, 128, 0, st>>>(a, ci, si, dist, dc, ds, n, max_steps, iters, rEH, e, dr);
    else
      k_bisect_band<32, 8><<<(unsigned)((n + 7) / 8), 256, 0, st>>>(a, ci, si, dist, dc, ds, n, max_steps, iters, rEH,
                                                               e, dr);
    return hipGetLastError();
  }
  // K halvings per round with 2^K lanes per angle: the fewest rounds with at
  // most about 65 536 lanes at once, then the halvings spread evenly over
  // those rounds (fewer lanes for the same number of rounds; the slowest
  // integration of a round sets its time).  K = 1: one lane per angle.
  int Kmax = 1;
  while (Kmax < 10 && Kmax < iters && n * (2L << Kmax) <= 65536) ++Kmax;
  const int rounds = (iters + Kmax - 1) / Kmax;
  const int K = rounds > 0 ? (iters + rounds - 1) / rounds : 1;
  if (K == 1 || iters <= 1) {
    k_bisect<<<grid_for(n, 64), 64, 0, st>>>(a, ci, si, dist, dc, ds, n, max_steps, iters, horizon(a), dr);
  } else {
    // one angle per block (2^K lanes, two or more waves) while the blocks
    // still fit one per CU: fewer waves per CU run the FP64 walk faster
    // (64 angles 8.9 -> 7.0 ms, 256 angles 8.6 -> 7.6 ms); beyond that,
    // blocks of at least 256 lanes (512 angles: 8.4 against 9.6 ms)
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      cus = 0;
    const bool one = (1 << K) >= 64 && n <= cus;
    const int block = one ? (1 << K) : ((1 << K) >= 256 ? (1 << K) : 256), G = block >> K;
    k_bisect_tree<<<(unsigned)((n + G - 1) / G), block, 0, st>>>(a, ci, si, dist, dc, ds, n, max_steps, iters,
                                                                  horizon(a), K, dr);
  }
  return hipGetLastError();
}

}  // extern "C"

namespace mk {
hipError_t shadow_bisect_on_stream(double a, double inc, const double* h_angles, long n, int max_steps,
                                   double error_allowed, int max_it, double dist, double* d_stage, double* h_radii,
                                   hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const int iters = bisect_iterations(error_allowed, max_it);
  std::vector<double> hc(n), hs(n), hr(n);
  for (long p = 0; p < n; ++p) {
    hc[p] = std::cos(h_angles[p]);
    hs[p] = std::sin(h_angles[p]);
  }
  predicted_shadow_radii(a, inc, h_angles, n, hr.data());
  double *dc = d_stage, *ds = d_stage + n, *dr = d_stage + 2 * n;
  hipError_t e = hipMemcpyAsync(dc, hc.data(), n * sizeof(double), hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(ds, hs.data(), n * sizeof(double), hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = launch_bisect(a, inc, dc, ds, n, max_steps, iters, dist, dr, hr.data(), st);
  if (e == hipSuccess) e = hipMemcpyAsync(h_radii, dr, n * sizeof(double), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  return e;
}
}  // namespace mk

extern "C" {

int mk_shadow_bisect(double a, double inc, const double* h_angles, long n, int max_steps, double error_allowed,
                     int max_it, double dist, double* h_radii) {
  MK_REQUIRE(n >= 0 && max_steps >= 1 && h_angles && h_radii, "mk_shadow_bisect: bad arguments");
  if (!n) return MK_OK;
  // every angle shares one bracket width, so the count is fixed up front
  const int iters = bisect_iterations(error_allowed, max_it);
  std::vector<double> hc(n), hs(n), hr(n);
  for (long p = 0; p < n; ++p) {
    hc[p] = std::cos(h_angles[p]);
    hs[p] = std::sin(h_angles[p]);
  }
  predicted_shadow_radii(a, inc, h_angles, n, hr.data());
  double *dc = nullptr, *ds = nullptr, *dr = nullptr;
  hipError_t e = hipMalloc(&dc, n * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&ds, n * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&dr, n * sizeof(double));
  if (e == hipSuccess) e = hipMemcpy(dc, hc.data(), n * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(ds, hs.data(), n * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = launch_bisect(a, inc, dc, ds, n, max_steps, iters, dist, dr, hr.data(), nullptr);
  if (e == hipSuccess) e = hipMemcpy(h_radii, dr, n * sizeof(double), hipMemcpyDeviceToHost);
  (void)hipFree(dc);
  (void)hipFree(ds);
  (void)hipFree(dr);
  if (e != hipSuccess) {
    set_error("mk_shadow_bisect: %s", hipGetErrorString(e));
    return MK_ERR_HIP;
  }
  return MK_OK;
}

int mk_shadow_bisect_noalloc(double a, double inc, const double* h_angles, long n, int max_steps,
                             double error_allowed, int max_it, double dist, double* h_radii) {
  MK_REQUIRE(n >= 0 && max_steps >= 1 && h_angles && h_radii, "mk_shadow_bisect_noalloc: bad arguments");
  MK_REQUIRE(n <= MK_NOALLOC_MAX_ANGLES, "mk_shadow_bisect_noalloc: %ld angles exceed the limit of %d", n,
             (int)MK_NOALLOC_MAX_ANGLES);
  if (!n) return MK_OK;
  const int iters = bisect_iterations(error_allowed, max_it);
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  double* scr = nullptr;
  int rc = noalloc_scratch(&scr);
  if (rc) return rc;
  double* hc = g_noalloc_host;
  double* hs = g_noalloc_host + MK_NOALLOC_MAX_ANGLES;
  for (long p = 0; p < n; ++p) {
    hc[p] = std::cos(h_angles[p]);
    hs[p] = std::sin(h_angles[p]);
  }
  // scratch: cos[0, MAX), sin[MAX, 2 MAX), radii[2 MAX, 3 MAX); every value
  // a launch reads (p < n) is written first, so nothing needs clearing
  double* dc = scr;
  double* ds = scr + MK_NOALLOC_MAX_ANGLES;
  double* dr = scr + 2 * MK_NOALLOC_MAX_ANGLES;
  hipError_t e = hipMemcpy(dc, hc, n * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(ds, hs, n * sizeof(double), hipMemcpyHostToDevice);
  // no prediction (building it allocates host memory): tree rounds, the
  // same radii
  if (e == hipSuccess) e = launch_bisect(a, inc, dc, ds, n, max_steps, iters, dist, dr, nullptr, nullptr);
  if (e == hipSuccess) e = hipMemcpy(h_radii, dr, n * sizeof(double), hipMemcpyDeviceToHost);
  if (e != hipSuccess) {
    set_error("mk_shadow_bisect_noalloc: %s", hipGetErrorString(e));
    return MK_ERR_HIP;
  }
  return MK_OK;
}

int mk_synchrotron(const double* Ne, const double* Th, const double* B, const double* pitch, const double* nu, long n,
                   int invariant, double rescale, double* j, double* al, void* s) {
  MK_REQUIRE(n >= 0 && Ne && Th && B && pitch && nu && j && al, "mk_synchrotron: bad arguments");
  if (!n) return MK_OK;
  k_synchrotron<<<grid_for(n, kBlock), kBlock, 0, as_stream(s)>>>(Ne, Th, B, pitch, nu, n, invariant, rescale, j, al);
  MK_LAUNCH_CHECK();
  return MK_OK;
}

int mk_solve_intensity(const double* j, const double* al, const double* dt, long T, long npx, double L, double* I,
                       double* dI, void* s) {
  MK_REQUIRE(T >= 0 && npx >= 0 && j && al && dt && I, "mk_solve_intensity: bad arguments");
  if (!npx) return MK_OK;
  k_solve<<<grid_for(npx, kBlock), kBlock, 0, as_stream(s)>>>(j, al, dt, T, npx, L, I, dI);
  MK_LAUNCH_CHECK();
  return MK_OK;
}

int mk_attenuated_emissivity(const double* j, const double* al, const double* dt, long T, long npx, double L,
                             double* out, void* s) {
  MK_REQUIRE(T >= 0 && npx >= 0 && j && al && dt && out, "mk_attenuated_emissivity: bad arguments");
  if (!npx || T < 2) return MK_OK;
  k_attenuated<<<grid_for(npx, kBlock), kBlock, 0, as_stream(s)>>>(j, al, dt, T, npx, L, out);
  MK_LAUNCH_CHECK();
  return MK_OK;
}

int mk_rlow_rhigh(const double* dens, const double* u, const double* beta, long n, double rl, double rh, double ge,
                  double gi, double* out, void* s) {
  MK_REQUIRE(n >= 0 && dens && u && beta && out, "mk_rlow_rhigh: bad arguments");
  if (!n) return MK_OK;
  k_rlow_rhigh<<<grid_for(n, kBlock), kBlock, 0, as_stream(s)>>>(dens, u, beta, n, rl, rh, ge, gi, out);
  MK_LAUNCH_CHECK();
  return MK_OK;
}

}  // extern "C"
