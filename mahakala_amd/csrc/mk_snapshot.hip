// mk_snapshot.hip -- device-resident AthenaK-shaped snapshots: the ghost-
// padded layout with its AMR halo fill, the leaf lookup grid, and the
// sampler kernels (grmhd/athenak.py:67-812 of the reference).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <map>
#include <vector>

#include "mk_device.h"
#include "mk_internal.h"
#include "mk_snapshot.h"

namespace mk {

// Neighbour record per (block, direction): type, up to 4 block ids, and the
// parity (t + d) mod 2 per axis used by coarse-neighbour injection.
enum NbType : int32_t { kNone = 0, kSame = 1, kCoarser = 2, kFiner = 3 };
constexpr int kNbRec = 8;  // type, id0..id3, odd_i, odd_j, odd_k

__device__ __forceinline__ int halo_dir(int I, int n) { return I == 0 ? -1 : (I == n + 1 ? 1 : 0); }

// One thread per padded cell of every block; writes the 8-prim record.
// Rules (athenak.py:212-514):
//   same level   copy of the facing cells of the neighbour
//   coarser      injection from the coarse cell that contains the halo cell
//   finer        mean of the 2x2x2 fine cells (faces, corners); the edge
//                branch reads, along the edge, the selector left by the
//                previous (v1, v2, v3) iteration of the reference's loop
//                (fine cell 0 on the first one), which this reproduces.
__global__ void __launch_bounds__(256) k_ghost_fill(const float* __restrict__ src, long nmb, int n,
                                                    const int32_t* __restrict__ var_index,
                                                    const int32_t* __restrict__ nb, float4* __restrict__ cells) {
  const int n2 = n + 2;
  const long cells_per_block = (long)n2 * n2 * n2;
  const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= nmb * cells_per_block) return;
  const long b = gid / cells_per_block;
  const int rem = (int)(gid - b * cells_per_block);
  const int K = rem / (n2 * n2), J = (rem / n2) % n2, I = rem % n2;
  const long nnn = (long)n * n * n;
  // file layout (prim, block, k, j, i)
  auto at = [&](int prim, long blk, int k, int j, int i) -> float {
    return src[((long)prim * nmb + blk) * nnn + ((long)k * n + j) * n + i];
  };
  float outv[8];
  const int di = halo_dir(I, n), dj = halo_dir(J, n), dk = halo_dir(K, n);
  if (di == 0 && dj == 0 && dk == 0) {
#pragma unroll
    for (int c = 0; c < 8; ++c) outv[c] = at(var_index[c], b, K - 1, J - 1, I - 1);
  } else {
    const int dir = (dk + 1) * 9 + (dj + 1) * 3 + (di + 1);
    const int32_t* rec = nb + ((long)b * 27 + dir) * kNbRec;
    const int type = rec[0];
    const int d[3] = {di, dj, dk};
    const int P[3] = {I, J, K};
    if (type == kNone) {
#pragma unroll
      for (int c = 0; c < 8; ++c) outv[c] = 0.0f;
    } else if (type == kSame) {
      int s[3];
#pragma unroll
      for (int ax = 0; ax < 3; ++ax) s[ax] = d[ax] == 1 ? 0 : (d[ax] == -1 ? n - 1 : P[ax] - 1);
#pragma unroll
      for (int c = 0; c < 8; ++c) outv[c] = at(var_index[c], rec[1], s[2], s[1], s[0]);
    } else if (type == kCoarser) {
      int s[3];
      const int h = n / 2;
#pragma unroll
      for (int ax = 0; ax < 3; ++ax) {
        const int odd = rec[5 + ax];
        if (d[ax] == 1) s[ax] = odd ? h : 0;
        else if (d[ax] == -1) s[ax] = odd ? n - 1 : h - 1;
        else s[ax] = (odd ? h : 0) + (P[ax] - 1) / 2;
      }
#pragma unroll
      for (int c = 0; c < 8; ++c) outv[c] = at(var_index[c], rec[1], s[2], s[1], s[0]);
    } else {
      const int h = n / 2;
      int nsl = 0, A = -1, e[3] = {0, 0, 0}, m[3] = {0, 0, 0}, src0[3];
#pragma unroll
      for (int ax = 0; ax < 3; ++ax) {
        src0[ax] = d[ax] == -1 ? n - 2 : 0;
        if (d[ax] == 0) {
          ++nsl;
          A = (A < 0) ? ax : A;
          e[ax] = (P[ax] - 1) / h;
          m[ax] = (P[ax] - 1) - e[ax] * h;
        }
      }
      // child block: corners one, edges indexed by e along the edge, faces
      // by e of the lower sliced axis + 2 e of the higher one
      int child;
      if (nsl == 0) child = rec[1];
      else if (nsl == 1) child = rec[1 + e[A]];
      else {
        int lo = -1, hi = -1;
        for (int ax = 0; ax < 3; ++ax)
          if (d[ax] == 0) { if (lo < 0) lo = ax; else hi = ax; }
        child = rec[1 + e[lo] + 2 * e[hi]];
      }
      double acc[8];
#pragma unroll
      for (int c = 0; c < 8; ++c) acc[c] = 0.0;
      int it = 0;
      for (int v1 = 0; v1 < 2; ++v1)
        for (int v2 = 0; v2 < 2; ++v2)
          for (int v3 = 0; v3 < 2; ++v3, ++it) {
            const int v[3] = {v1, v2, v3};
            int s[3];
#pragma unroll
            for (int ax = 0; ax < 3; ++ax) {
              if (d[ax] != 0) {
                s[ax] = src0[ax] + v[ax];
              } else if (nsl == 1) {
                if (it == 0) {
                  s[ax] = 0;
                } else {
                  const int pv = (ax == 0) ? ((it - 1) >> 2) & 1 : ((ax == 1) ? ((it - 1) >> 1) & 1 : (it - 1) & 1);
                  s[ax] = 2 * m[ax] + pv;
                }
              } else {
                s[ax] = 2 * m[ax] + v[ax];
              }
            }
#pragma unroll
            for (int c = 0; c < 8; ++c) acc[c] += (double)(at(var_index[c], child, s[2], s[1], s[0]) / 8.0f);
          }
#pragma unroll
      for (int c = 0; c < 8; ++c) outv[c] = (float)acc[c];
    }
  }
  cells[2 * gid] = make_float4(outv[0], outv[1], outv[2], outv[3]);
  cells[2 * gid + 1] = make_float4(outv[4], outv[5], outv[6], outv[7]);
}

__global__ void k_block_index(SnapView sv, const double* S, long n, int stride, int32_t* mb) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const double* q = S + (long)stride * p;
  mb[p] = find_block(sv, sv.meta, sv.grid, q[1], q[2], q[3]);
}

__global__ void __launch_bounds__(256) k_sample_prims(SnapView sv, const double* S, long n, double* out) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const double* q = S + 8 * p;
  double P[8];
  trilinear<double>(sv, q[1], q[2], q[3], P);
#pragma unroll
  for (int c = 0; c < 8; ++c) out[c * n + p] = P[c];
}

__global__ void __launch_bounds__(256) k_sample_scalars(SnapView sv, const double* S, long n, double a, double fbc,
                                                        double* out) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  double q[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) q[k] = S[8 * p + k];
  double P[8], sc[5];
  trilinear<double>(sv, q[1], q[2], q[3], P);
  fluid_scalars<double>(P, q[1], q[2], q[3], q[4], q[5], q[6], q[7], a, fbc, sc);
  sc[2] = ::acos(sc[2]);  // pitch angle from its cosine
#pragma unroll
  for (int c = 0; c < 5; ++c) out[c * n + p] = sc[c];
}

static long pymod(long a, long m) { long r = a % m; return r < 0 ? r + m : r; }
static long pyfloordiv(long a, long m) { return (a - pymod(a, m)) / m; }

}  // namespace mk

using namespace mk;

mk::SnapView mk_snapshot::view() const {
  SnapView v;
  v.cells = d_cells;
  v.meta = d_meta;
  v.grid = d_grid;
  v.nmb = (int)nmb;
  v.n2 = n + 2;
  v.gx = G[0];
  v.gy = G[1];
  v.gz = G[2];
  for (int d = 0; d < 3; ++d) {
    v.g0[d] = g0[d];
    v.ginv[d] = ginv[d];
    v.dom_lo[d] = dom_lo[d];
    v.dom_hi[d] = dom_hi[d];
  }
  return v;
}

extern "C" {

int mk_snapshot_create(const float* h_prims, const int32_t* h_var, long nmb, int n, const int32_t* h_levels,
                       const int32_t* h_logical, const double* x1f, const double* x2f, const double* x3f,
                       const double* x1v, const double* x2v, const double* x3v, mk_snapshot** out) {
  MK_REQUIRE(out && h_prims && h_var && h_levels && h_logical && x1f && x2f && x3f && x1v && x2v && x3v,
             "mk_snapshot_create: null argument");
  MK_REQUIRE(nmb >= 1 && n >= 2 && n % 2 == 0, "mk_snapshot_create: need nmb >= 1 and an even block size n >= 2");
  for (int c = 0; c < 8; ++c) MK_REQUIRE(h_var[c] >= 0 && h_var[c] < 8, "mk_snapshot_create: bad variable index");
  *out = nullptr;
  // (level, i, j, k) -> block id; the reference's dict (athenak.py:108-113)
  std::map<std::array<long, 4>, long> index;
  for (long b = 0; b < nmb; ++b) {
    std::array<long, 4> key = {h_levels[b], h_logical[3 * b], h_logical[3 * b + 1], h_logical[3 * b + 2]};
    if (index.count(key)) {
      set_error("mk_snapshot_create: duplicate meshblock (level %ld, %ld, %ld, %ld)", key[0], key[1], key[2], key[3]);
      return MK_ERR_MESH;
    }
    index[key] = b;
  }
  auto find = [&](long l, long i, long j, long k) -> long {
    auto it = index.find({l, i, j, k});
    return it == index.end() ? -1 : it->second;
  };
  std::vector<int32_t> nb((size_t)nmb * 27 * kNbRec, 0);
  for (long b = 0; b < nmb; ++b) {
    const long L = h_levels[b], t[3] = {h_logical[3 * b], h_logical[3 * b + 1], h_logical[3 * b + 2]};
    for (int dk = -1; dk <= 1; ++dk)
      for (int dj = -1; dj <= 1; ++dj)
        for (int di = -1; di <= 1; ++di) {
          if (!di && !dj && !dk) continue;
          const int d[3] = {di, dj, dk};
          int32_t* rec = &nb[((size_t)b * 27 + (dk + 1) * 9 + (dj + 1) * 3 + (di + 1)) * kNbRec];
          long id = find(L, t[0] + di, t[1] + dj, t[2] + dk);
          if (id >= 0) {
            rec[0] = kSame;
            rec[1] = (int32_t)id;
            continue;
          }
          id = find(L - 1, pyfloordiv(t[0] + di, 2), pyfloordiv(t[1] + dj, 2), pyfloordiv(t[2] + dk, 2));
          if (id >= 0) {
            rec[0] = kCoarser;
            rec[1] = (int32_t)id;
            for (int ax = 0; ax < 3; ++ax) rec[5 + ax] = (int32_t)pymod(t[ax] + d[ax], 2);
            continue;
          }
          long base[3];
          for (int ax = 0; ax < 3; ++ax) base[ax] = 2 * (t[ax] + d[ax]);
          if (find(L + 1, base[0], base[1], base[2]) < 0) continue;  // domain edge: halo stays 0
          for (int ax = 0; ax < 3; ++ax)
            if (d[ax] == -1) base[ax] += 1;
          int sl[3], nsl = 0;
          for (int ax = 0; ax < 3; ++ax)
            if (d[ax] == 0) sl[nsl++] = ax;
          rec[0] = kFiner;
          const int nchild = 1 << nsl;
          for (int c = 0; c < nchild; ++c) {
            long k[3] = {base[0], base[1], base[2]};
            for (int q = 0; q < nsl; ++q) k[sl[q]] += (c >> q) & 1;
            const long cid = find(L + 1, k[0], k[1], k[2]);
            if (cid < 0) {
              set_error("mk_snapshot_create: missing finer neighbour of block %ld (level %ld)", b, L);
              return MK_ERR_MESH;
            }
            rec[1 + c] = (int32_t)cid;
          }
        }
  }
  // per-block geometry: lo, hi (faces), q0 = v0 - dx, 1 / dx (dx = v1 - v0)
  std::vector<double> meta((size_t)nmb * kMeta, 0.0);
  const double* f[3] = {x1f, x2f, x3f};
  const double* v[3] = {x1v, x2v, x3v};
  for (long b = 0; b < nmb; ++b)
    for (int ax = 0; ax < 3; ++ax) {
      const double dx = v[ax][b * n + 1] - v[ax][b * n];
      if (!(dx > 0)) {
        set_error("mk_snapshot_create: block %ld has non-increasing cell centres", b);
        return MK_ERR_MESH;
      }
      meta[kMeta * b + ax] = f[ax][b * (n + 1)];
      meta[kMeta * b + 3 + ax] = f[ax][b * (n + 1) + n];
      meta[kMeta * b + 6 + ax] = v[ax][b * n] - dx;
      meta[kMeta * b + 9 + ax] = 1.0 / dx;
    }
  // lookup grid at the finest level's block size
  mk_snapshot* s = new mk_snapshot();
  s->nmb = nmb;
  s->n = n;
  long Lmax = h_levels[0];
  long bmax = 0;
  for (long b = 0; b < nmb; ++b)
    if (h_levels[b] > Lmax) { Lmax = h_levels[b]; bmax = b; }
  for (int ax = 0; ax < 3; ++ax) {
    double lo = meta[ax], hi = meta[3 + ax];
    for (long b = 0; b < nmb; ++b) {
      lo = std::min(lo, meta[kMeta * b + ax]);
      hi = std::max(hi, meta[kMeta * b + 3 + ax]);
    }
    const double w = meta[kMeta * bmax + 3 + ax] - meta[kMeta * bmax + ax];
    if (!(w > 0)) {
      delete s;
      set_error("mk_snapshot_create: degenerate block extent");
      return MK_ERR_MESH;
    }
    s->dom_lo[ax] = lo;
    s->dom_hi[ax] = hi;
    s->g0[ax] = lo;
    s->ginv[ax] = 1.0 / w;
    s->G[ax] = (int)std::llround((hi - lo) / w);
    if (s->G[ax] < 1 || s->G[ax] > 4096) {
      delete s;
      set_error("mk_snapshot_create: lookup grid of %d cells per axis is out of range", s->G[ax]);
      return MK_ERR_MESH;
    }
  }
  const size_t gcells = (size_t)s->G[0] * s->G[1] * s->G[2];
  std::vector<int32_t> grid(gcells, -1);
  for (long b = 0; b < nmb; ++b) {
    long c0[3], c1[3];
    for (int ax = 0; ax < 3; ++ax) {
      c0[ax] = std::llround((meta[kMeta * b + ax] - s->g0[ax]) * s->ginv[ax]);
      c1[ax] = std::llround((meta[kMeta * b + 3 + ax] - s->g0[ax]) * s->ginv[ax]);
      c0[ax] = std::max(0L, c0[ax]);
      c1[ax] = std::min((long)s->G[ax], c1[ax]);
    }
    for (long z = c0[2]; z < c1[2]; ++z)
      for (long y = c0[1]; y < c1[1]; ++y)
        for (long x = c0[0]; x < c1[0]; ++x) grid[((size_t)z * s->G[1] + y) * s->G[0] + x] = (int32_t)b;
  }
  // device buffers
  const long n2 = n + 2;
  const size_t ncells = (size_t)nmb * n2 * n2 * n2;
  const size_t src_bytes = (size_t)8 * nmb * n * n * n * sizeof(float);
  float* d_src = nullptr;
  int32_t *d_nb = nullptr, *d_var = nullptr;
  hipError_t e = hipMalloc(&s->d_cells, ncells * 2 * sizeof(float4));
  if (e == hipSuccess) e = hipMalloc(&s->d_meta, meta.size() * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&s->d_grid, gcells * sizeof(int32_t));
  if (e == hipSuccess) e = hipMalloc(&d_src, src_bytes);
  if (e == hipSuccess) e = hipMalloc(&d_nb, nb.size() * sizeof(int32_t));
  if (e == hipSuccess) e = hipMalloc(&d_var, 8 * sizeof(int32_t));
  if (e == hipSuccess) e = hipMemcpy(s->d_meta, meta.data(), meta.size() * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(s->d_grid, grid.data(), gcells * sizeof(int32_t), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d_src, h_prims, src_bytes, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d_nb, nb.data(), nb.size() * sizeof(int32_t), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d_var, h_var, 8 * sizeof(int32_t), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    k_ghost_fill<<<grid_for((long)ncells, 256), 256>>>(d_src, nmb, n, d_var, d_nb, s->d_cells);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  (void)hipFree(d_src);
  (void)hipFree(d_nb);
  (void)hipFree(d_var);
  if (e != hipSuccess) {
    set_error("mk_snapshot_create: %s", hipGetErrorString(e));
    mk_snapshot_free(s);
    return e == hipErrorOutOfMemory ? MK_ERR_NOMEM : MK_ERR_HIP;
  }
  s->bytes = ncells * 2 * sizeof(float4) + meta.size() * sizeof(double) + gcells * sizeof(int32_t);
  *out = s;
  return MK_OK;
}

int mk_snapshot_free(mk_snapshot* s) {
  if (!s) return MK_OK;
  (void)hipFree(s->d_cells);
  (void)hipFree(s->d_meta);
  (void)hipFree(s->d_grid);
  delete s;
  return MK_OK;
}

int mk_snapshot_bytes(const mk_snapshot* s, size_t* bytes) {
  MK_REQUIRE(s && bytes, "mk_snapshot_bytes: null argument");
  *bytes = s->bytes;
  return MK_OK;
}

int mk_snapshot_download(const mk_snapshot* s, float* h) {
  MK_REQUIRE(s && h, "mk_snapshot_download: null argument");
  const size_t n2 = s->n + 2;
  MK_HIP(hipMemcpy(h, s->d_cells, (size_t)s->nmb * n2 * n2 * n2 * 8 * sizeof(float), hipMemcpyDeviceToHost));
  return MK_OK;
}

int mk_block_index(const mk_snapshot* s, const double* d_S, long n, int stride, int32_t* d_mb, void* st) {
  MK_REQUIRE(s && d_S && d_mb && n >= 0 && stride >= 4, "mk_block_index: bad arguments");
  if (!n) return MK_OK;
  k_block_index<<<grid_for(n, 256), 256, 0, as_stream(st)>>>(s->view(), d_S, n, stride, d_mb);
  MK_LAUNCH_CHECK();
  return MK_OK;
}

int mk_sample_prims(const mk_snapshot* s, const double* d_S, long n, double* d_out, void* st) {
  MK_REQUIRE(s && d_S && d_out && n >= 0, "mk_sample_prims: bad arguments");
  if (!n) return MK_OK;
  k_sample_prims<<<grid_for(n, 256), 256, 0, as_stream(st)>>>(s->view(), d_S, n, d_out);
  MK_LAUNCH_CHECK();
  return MK_OK;
}

int mk_sample_scalars(const mk_snapshot* s, const double* d_S, long n, double a, double fallback, double* d_out,
                      void* st) {
  MK_REQUIRE(s && d_S && d_out && n >= 0, "mk_sample_scalars: bad arguments");
  if (!n) return MK_OK;
  k_sample_scalars<<<grid_for(n, 256), 256, 0, as_stream(st)>>>(s->view(), d_S, n, a, std::cos(fallback), d_out);
  MK_LAUNCH_CHECK();
  return MK_OK;
}

}  // extern "C"
