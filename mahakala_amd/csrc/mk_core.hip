// mk_core.hip -- runtime helpers, camera, metric, integrator, shadow
// bisection and transfer kernels of the per-ray hot path (gfx950).
#include <hip/hip_runtime.h>

#include <cmath>
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
  k_integrate_final<<<grid_for(n, kBlock), kBlock, 0, as_stream(s)>>>(d_s0, n, N, div, tol, a, horizon(a), d_K,
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
  k_integrate_history<<<grid_for(n, kBlock), kBlock, 0, as_stream(s)>>>(d_s0, n, div, tol, a, horizon(a), T, d_S,
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
    k_select<<<grid_for(n, kBlock), kBlock>>>(a, ci, si, dist, dx, dy, n, max_steps, horizon(a), dout);
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

// Launches the bisection of n angles (cos and sin in dc, ds; device) into dr
// on the null stream.
static hipError_t launch_bisect(double a, double inc, const double* dc, const double* ds, long n, int max_steps,
                                int iters, double dist, double* dr) {
  double ci, si;
  cam_angles(inc, &ci, &si);
  // K halvings per round with 2^K lanes per angle: the fewest rounds with at
  // most about 65 536 lanes at once, then the halvings spread evenly over
  // those rounds (fewer lanes for the same number of rounds; the slowest
  // integration of a round sets its time).  K = 1: one lane per angle.
  int Kmax = 1;
  while (Kmax < 10 && Kmax < iters && n * (2L << Kmax) <= 65536) ++Kmax;
  const int rounds = (iters + Kmax - 1) / Kmax;
  const int K = rounds > 0 ? (iters + rounds - 1) / rounds : 1;
  if (K == 1 || iters <= 1) {
    k_bisect<<<grid_for(n, 64), 64>>>(a, ci, si, dist, dc, ds, n, max_steps, iters, horizon(a), dr);
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
    k_bisect_tree<<<(unsigned)((n + G - 1) / G), block>>>(a, ci, si, dist, dc, ds, n, max_steps, iters, horizon(a),
                                                           K, dr);
  }
  return hipGetLastError();
}

int mk_shadow_bisect(double a, double inc, const double* h_angles, long n, int max_steps, double error_allowed,
                     int max_it, double dist, double* h_radii) {
  MK_REQUIRE(n >= 0 && max_steps >= 1 && h_angles && h_radii, "mk_shadow_bisect: bad arguments");
  if (!n) return MK_OK;
  // every angle shares one bracket width, so the count is fixed up front
  const int iters = bisect_iterations(error_allowed, max_it);
  std::vector<double> hc(n), hs(n);
  for (long p = 0; p < n; ++p) {
    hc[p] = std::cos(h_angles[p]);
    hs[p] = std::sin(h_angles[p]);
  }
  double *dc = nullptr, *ds = nullptr, *dr = nullptr;
  hipError_t e = hipMalloc(&dc, n * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&ds, n * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&dr, n * sizeof(double));
  if (e == hipSuccess) e = hipMemcpy(dc, hc.data(), n * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(ds, hs.data(), n * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = launch_bisect(a, inc, dc, ds, n, max_steps, iters, dist, dr);
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
  if (e == hipSuccess) e = launch_bisect(a, inc, dc, ds, n, max_steps, iters, dist, dr);
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
