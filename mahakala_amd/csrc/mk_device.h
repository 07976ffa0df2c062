// mk_device.h -- device-side physics of the per-ray hot path (gfx950).
//
// Everything here runs in registers of one lane per ray.  The arithmetic
// follows the reference (file:line in mahakala/):
//   Kerr-Schild metric            geodesics.py:88-104
//   geodesic right-hand side      geodesics.py:294-309  (jacfwd + LU there;
//                                 here the closed-form gradient of the same
//                                 expression and g^{mu nu} = eta - f l^mu l^nu)
//   RK4                           geodesics.py:317-336
//   radius, step rule             geodesics.py:249-252, 284-291
//   camera + nullify              geodesics.py:58-85, 107-134, 204-216
//   sampler                       grmhd/athenak.py:712-794
//   electrons, synchrotron        electrons.py:32-50, transfer.py:30-86
//   per-sample glue               images.py:86-118
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mk {

// constants.py:23-31 (electrons.py and grmhd.py repeat MP)
constexpr double kEE = 4.8032e-10;
constexpr double kCL = 2.99792458e10;
constexpr double kME = 9.1094e-28;
constexpr double kMP = 1.6726e-24;
constexpr double kHPL = 6.6261e-27;
constexpr double kPI = 3.141592653589793;

// ---------------------------------------------------------------- precision
// FP64: division and square root on the geodesic right-hand side use the
// hardware reciprocal / reciprocal square root refined by Newton and
// Goldschmidt steps (within 2 ulp of IEEE; positive normal arguments only
// there); the step rule, radius and camera keep IEEE operations.  FP32: hardware reciprocal,
// square root, exp2 and log2 (about 1 ulp).
template <typename T> struct Math;
template <> struct Math<double> {
  static __device__ __forceinline__ double rcp(double x) {
    double y = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, y, 1.0);
    y = __builtin_fma(y, e, y);
    e = __builtin_fma(-x, y, 1.0);
    return __builtin_fma(y, e, y);
  }
  // 1/x and 1/y from one refined reciprocal of x y (x y finite and normal;
  // each within 1.9 ulp where separate reciprocals round correctly): one
  // quarter-rate estimate and four FMAs fewer
  static __device__ __forceinline__ void rcp2(double x, double y, double& ix, double& iy) {
    const double q = rcp(x * y);
    ix = y * q;
    iy = x * q;
  }
  static __device__ __forceinline__ double div(double a, double b) { return a / b; }
  static __device__ __forceinline__ double sqrt(double x) {
    // Goldschmidt refinement of rsq, then two residual corrections
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = 0.5 * y;
    const double r = __builtin_fma(-g, h, 0.5);
    g = __builtin_fma(g, r, g);
    h = __builtin_fma(h, r, h);
    double d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    return x == 0.0 ? x : g;
  }
  static __device__ __forceinline__ double sqrt_ieee(double x) { return __builtin_sqrt(x); }
  // sqrt(x) and 1/sqrt(x) from one reciprocal square root: two Goldschmidt
  // steps carry g -> sqrt(x) and h -> 1/(2 sqrt(x)) from the estimate's
  // 2^-24 to about 1.5 ulp, one residual correction rounds g as sqrt()
  // does (the same bits), and 1/sqrt(x) = 2h (within 1.5 ulp, as a refined
  // reciprocal of sqrt(x) is) without a second quarter-rate estimate.  For
  // x = 0 (on the right-hand side only at the ring singularity, r = 0) s is
  // NaN, not 0; the right-hand side is NaN there either way (is = inf)
  static __device__ __forceinline__ void sqrt_rcp(double x, double& s, double& is) {
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = 0.5 * y;
    double r = __builtin_fma(-g, h, 0.5);
    g = __builtin_fma(g, r, g);
    h = __builtin_fma(h, r, h);
    r = __builtin_fma(-g, h, 0.5);
    g = __builtin_fma(g, r, g);
    h = __builtin_fma(h, r, h);
    const double d = __builtin_fma(-g, g, x);
    s = __builtin_fma(d, h, g);
    is = 2.0 * h;
  }
  static __device__ __forceinline__ double exp(double x) { return ::exp(x); }
  static __device__ __forceinline__ double pow(double x, double y) { return ::pow(x, y); }
  static __device__ __forceinline__ double cbrt(double x) { return ::cbrt(x); }
  static __device__ __forceinline__ double sin(double x) { return ::sin(x); }
  static __device__ __forceinline__ double acos(double x) { return ::acos(x); }
  static __device__ __forceinline__ double cos(double x) { return ::cos(x); }
  static __device__ __forceinline__ double abs(double x) { return __builtin_fabs(x); }
};
template <> struct Math<float> {
  static __device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
  // separate hardware reciprocals: in FP32 the shared form would trade one
  // quarter-rate estimate for three multiplies, about the same issue time
  static __device__ __forceinline__ void rcp2(float x, float y, float& ix, float& iy) {
    ix = __builtin_amdgcn_rcpf(x);
    iy = __builtin_amdgcn_rcpf(y);
  }
  static __device__ __forceinline__ float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
  static __device__ __forceinline__ float sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
  static __device__ __forceinline__ float sqrt_ieee(float x) { return __builtin_amdgcn_sqrtf(x); }
  // sqrt(x) and 1/sqrt(x) from one hardware reciprocal square root
  static __device__ __forceinline__ void sqrt_rcp(float x, float& s, float& is) {
    is = __builtin_amdgcn_rsqf(x);
    s = x * is;
  }
  static __device__ __forceinline__ float exp(float x) { return __builtin_amdgcn_exp2f(x * 1.4426950408889634f); }
  static __device__ __forceinline__ float pow(float x, float y) {
    // x >= 0 here; log2(0) = -inf gives 0, log2(inf) = inf gives inf
    return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x));
  }
  static __device__ __forceinline__ float cbrt(float x) {
    return __builtin_amdgcn_exp2f(0.3333333333333333f * __builtin_amdgcn_logf(x));
  }
  static __device__ __forceinline__ float sin(float x) { return ::__sinf(x); }
  static __device__ __forceinline__ float acos(float x) { return ::acosf(x); }
  static __device__ __forceinline__ float cos(float x) { return ::__cosf(x); }
  static __device__ __forceinline__ float abs(float x) { return __builtin_fabsf(x); }
};

template <typename T> __device__ __forceinline__ bool is_nan(T x) { return x != x; }

// ------------------------------------------------------------------ metric
// g = eta + f l l, l_mu = (1, (r x + a y)/(r^2+a^2), (r y - a x)/(r^2+a^2), z/r)
template <typename T> struct KS {
  T f, l1, l2, l3;
};

template <typename T>
__device__ __forceinline__ KS<T> ks_metric(T x, T y, T z, T a) {
  using M = Math<T>;
  const T aa = a * a, zz = z * z;
  const T kk = T(0.5) * (x * x + y * y + zz - aa);
  const T rr = M::sqrt(kk * kk + aa * zz) + kk;
  const T r = M::sqrt(rr);
  KS<T> m;
  m.f = (T(2) * rr * r) * M::rcp(rr * rr + aa * zz);
  const T s = M::rcp(rr + aa);
  m.l1 = (r * x + a * y) * s;
  m.l2 = (r * y - a * x) * s;
  m.l3 = z * M::rcp(r);
  return m;
}

// dv/dlambda = -Gamma^alpha_{mu nu} v^mu v^nu for g = eta + f l l
// (geodesics.py:294-309: g^{ab} (-d_s g_{bn} v^s v^n + 1/2 d_b g_{mn} v^m v^n)).
// With L = l.v, W_i = v^s d_s l_i, D_s = d_s l_mu v^mu, F_s = d_s f:
//   c_b = 1/2 F_b L^2 + f L D_b - l_b (L F.v + f W.v) - f L W_b
//   a^alpha = eta^{alpha b} c_b - f l^alpha (l^b c_b),  l^mu = (-1, l_i).
// The Jacobian of l_i is d_j l_i = A_i d_j(r^2) + B_ij with
//   A = (s (x/2r - l1), s (y/2r - l2), -l3 / 2r^2),  s = 1/(r^2 + a^2),
//   B = [[s r, s a, 0], [-s a, s r, 0], [0, 0, 1/r]],
// so W, D and W.v reduce to two dot products (dv = d(r^2).v, Av = A.v):
//   D - W = d(r^2) Av - A dv + (-2 s a v2, 2 s a v1, 0),
//   W.v = dv Av + s r (v1^2 + v2^2) + v3^2 / r,
// and F = kappa d(r^2) - zeta e_z, kappa = f (3/2r^2 - 2 r^2/den),
// zeta = 2 f a^2 z / den, den = r^4 + a^2 z^2.
template <typename T>
__device__ __forceinline__ void ks_accel(T x, T y, T z, T v0, T v1, T v2, T v3, T a, T acc[4]) {
  using M = Math<T>;
  const T aa = a * a, zz = z * z;
  const T kk = T(0.5) * (x * x + y * y + zz - aa);
  T Q, iQ, r, ir;
  M::sqrt_rcp(kk * kk + aa * zz, Q, iQ);
  const T rr = Q + kk;
  M::sqrt_rcp(rr, r, ir);
  T iden, s;
  M::rcp2(rr * rr + aa * zz, rr + aa, iden, s);
  // products shared below: t2 = 2 r^2/den (f = t2 r, and kappa's second
  // term), rs = r s and as = a s (l1, l2, W.v, the rotation term of D - W),
  // hi = 1/(2 r^2) (A3, kappa); 2 a^2 is uniform
  const T t2 = T(2) * rr * iden;
  const T f = t2 * r;
  const T rs = r * s, as = a * s;
  const T l1 = rs * x + as * y;
  const T l2 = rs * y - as * x;
  const T l3 = z * ir;
  // gradient of r^2, and A
  const T g = rr * iQ;
  const T d1 = x * g, d2 = y * g, d3 = z * (rr + aa) * iQ;
  const T h = T(0.5) * ir, hi = h * ir;
  const T A1 = s * (h * x - l1), A2 = s * (h * y - l2), A3 = -l3 * hi;
  const T kap = f * (T(3) * hi - t2);
  const T zeta = f * iden * (T(2) * aa) * z;
  const T L = v0 + l1 * v1 + l2 * v2 + l3 * v3;
  const T dv = d1 * v1 + d2 * v2 + d3 * v3;
  const T Av = A1 * v1 + A2 * v2 + A3 * v3;
  const T Wv = dv * Av + rs * (v1 * v1 + v2 * v2) + ir * (v3 * v3);
  const T P = L * (kap * dv - zeta * v3) + f * Wv;
  const T fL = f * L, hL2 = T(0.5) * L * L;
  const T mu = hL2 * kap + fL * Av, nu = fL * dv;
  const T sa2 = T(2) * as * fL;
  const T c1 = mu * d1 - nu * A1 - sa2 * v2 - l1 * P;
  const T c2 = mu * d2 - nu * A2 + sa2 * v1 - l2 * P;
  const T c3 = mu * d3 - nu * A3 - l3 * P - hL2 * zeta;
  const T flc = f * (P + l1 * c1 + l2 * c2 + l3 * c3);
  acc[0] = P + flc;
  acc[1] = c1 - l1 * flc;
  acc[2] = c2 - l2 * flc;
  acc[3] = c3 - l3 * flc;
}

// RK4 over the 8-vector state, all in T (geodesics.py:317-336).  The stage
// slopes stay unscaled: t = s + (h/2) F is one FMA, and h/6 scales the slope
// sum F1 + 2 F2 + 2 F3 + F4 once, which saves the 32 products k = h F of the
// scaled form (the roundings differ from it at the ulp level).
template <typename T>
__device__ __forceinline__ void rk4_step(const T s[8], T dt, T a, T out[8]) {
  const T hh = T(0.5) * dt, h6 = dt * T(1.0 / 6.0);
  T F[8], acc[4], t[8], sum[8];
  // stage 1
  ks_accel<T>(s[1], s[2], s[3], s[4], s[5], s[6], s[7], a, acc);
#pragma unroll
  for (int i = 0; i < 4; ++i) { F[i] = s[4 + i]; F[4 + i] = acc[i]; }
#pragma unroll
  for (int i = 0; i < 8; ++i) { sum[i] = F[i]; t[i] = s[i] + hh * F[i]; }
  // stage 2
  ks_accel<T>(t[1], t[2], t[3], t[4], t[5], t[6], t[7], a, acc);
#pragma unroll
  for (int i = 0; i < 4; ++i) { F[i] = t[4 + i]; F[4 + i] = acc[i]; }
#pragma unroll
  for (int i = 0; i < 8; ++i) { sum[i] += T(2) * F[i]; t[i] = s[i] + hh * F[i]; }
  // stage 3
  ks_accel<T>(t[1], t[2], t[3], t[4], t[5], t[6], t[7], a, acc);
#pragma unroll
  for (int i = 0; i < 4; ++i) { F[i] = t[4 + i]; F[4 + i] = acc[i]; }
#pragma unroll
  for (int i = 0; i < 8; ++i) { sum[i] += T(2) * F[i]; t[i] = s[i] + dt * F[i]; }
  // stage 4
  ks_accel<T>(t[1], t[2], t[3], t[4], t[5], t[6], t[7], a, acc);
#pragma unroll
  for (int i = 0; i < 4; ++i) { F[i] = t[4 + i]; F[4 + i] = acc[i]; }
#pragma unroll
  for (int i = 0; i < 8; ++i) out[i] = s[i] + h6 * (sum[i] + F[i]);
}

// Mixed step: stages in FP32, increment added to the FP64 state.  The stage
// slopes stay unscaled (t = s + (h/2) F is one FMA, the slope sum is
// F1 + 2 F2 + 2 F3 + F4 and h/6 scales it once), which saves the 32 products
// k = h F of the scaled form.  Coordinate time (component 0) enters no
// right-hand side and no image, so this step carries it over unchanged.
__device__ __forceinline__ void rk4_step_mixed(const double s[8], double dt, float a, double out[8]) {
  float sf[8];
#pragma unroll
  for (int i = 1; i < 8; ++i) sf[i] = (float)s[i];
  const float h = (float)dt, hh = 0.5f * h, h6 = h * (1.0f / 6.0f);
  float F[8], acc[4], t[8], sum[8];
  ks_accel<float>(sf[1], sf[2], sf[3], sf[4], sf[5], sf[6], sf[7], a, acc);
#pragma unroll
  for (int i = 1; i < 4; ++i) F[i] = sf[4 + i];
#pragma unroll
  for (int i = 0; i < 4; ++i) F[4 + i] = acc[i];
#pragma unroll
  for (int i = 1; i < 8; ++i) { sum[i] = F[i]; t[i] = sf[i] + hh * F[i]; }
  ks_accel<float>(t[1], t[2], t[3], t[4], t[5], t[6], t[7], a, acc);
#pragma unroll
  for (int i = 1; i < 4; ++i) F[i] = t[4 + i];
#pragma unroll
  for (int i = 0; i < 4; ++i) F[4 + i] = acc[i];
#pragma unroll
  for (int i = 1; i < 8; ++i) { sum[i] += 2.0f * F[i]; t[i] = sf[i] + hh * F[i]; }
  ks_accel<float>(t[1], t[2], t[3], t[4], t[5], t[6], t[7], a, acc);
#pragma unroll
  for (int i = 1; i < 4; ++i) F[i] = t[4 + i];
#pragma unroll
  for (int i = 0; i < 4; ++i) F[4 + i] = acc[i];
#pragma unroll
  for (int i = 1; i < 8; ++i) { sum[i] += 2.0f * F[i]; t[i] = sf[i] + h * F[i]; }
  ks_accel<float>(t[1], t[2], t[3], t[4], t[5], t[6], t[7], a, acc);
#pragma unroll
  for (int i = 1; i < 4; ++i) F[i] = t[4 + i];
#pragma unroll
  for (int i = 0; i < 4; ++i) F[4 + i] = acc[i];
  out[0] = s[0];
#pragma unroll
  for (int i = 1; i < 8; ++i) out[i] = s[i] + (double)(h6 * (sum[i] + F[i]));
}

// geodesics.py:284-291.  The square roots are IEEE ones: the compiler's
// FP64 sqrt is this same rsq estimate, Goldschmidt step and two residual
// corrections (Math<double>::sqrt), wrapped in a rescaling for arguments
// below 2^-767 and a pass-through for 0 and inf; the arguments here are 0 or
// far above 2^-767, and an infinite one freezes the ray either way.
__device__ __forceinline__ double radius_cal(double x, double y, double z, double a) {
#pragma clang fp contract(off)
  using M = Math<double>;
  const double R = M::sqrt(x * x + y * y + z * z);
  const double R2 = R * R, aa = a * a;
  return M::sqrt((R2 - aa + M::sqrt((R2 - aa) * (R2 - aa) + 4.0 * aa * (z * z))) / 2.0);
}

// radius_cal in FP32 from the FP64 position (the mixed step rule: FP32
// resolves |r - r_EH| to ~1e-7 M against the 1e-4 tolerance, and r to
// ~1e-4 M against the 1500 M escape bound)
__device__ __forceinline__ double radius_f32(double x, double y, double z, double a) {
  const float xf = (float)x, yf = (float)y, zf = (float)z, af = (float)a;
  const float R2 = xf * xf + yf * yf + zf * zf, aa = af * af;
  const float w = R2 - aa;
  return (double)__builtin_amdgcn_sqrtf(0.5f * (w + __builtin_amdgcn_sqrtf(w * w + 4.0f * aa * (zf * zf))));
}

// step rule with the division by div replaced by a product with 1/div
__device__ __forceinline__ double step_rule_fast(double r, double rEH, double inv_div, double div, double tol) {
  const double dt = -(r - rEH) * inv_div;
  const double m = __builtin_fabs(dt) * div;
  if (dt != dt || m < tol || m > 1500.0) return 0.0;
  return dt;
}

// geodesics.py:249-252: dt = -(r - r_EH)/div, zeroed on NaN, |dt| div < tol
// or |dt| div > 1500.  The quotient is the IEEE one: with y = 1/div rounded
// (loop-invariant, so computed once) and q = num y, the residual num - q div
// is exact in one FMA and q + residual y rounds to num/div (Markstein's
// theorem; no overflow or underflow here).  A zero num gives +0 where the
// division gives -0, and both are frozen below.
__device__ __forceinline__ double step_rule(double r, double rEH, double div, double tol) {
#pragma clang fp contract(off)
  const double num = -(r - rEH);
  const double y = 1.0 / div;
  const double q = num * y;
  const double dt = __builtin_fma(__builtin_fma(-q, div, num), y, q);
  const double m = __builtin_fabs(dt) * div;
  if (dt != dt || m < tol || m > 1500.0) return 0.0;
  return dt;
}

// ------------------------------------------------------------------ camera
// _Image_to_BH / _perpendicular / cross product (geodesics.py:204-216,
// 137-181, 107-134), evaluated without contraction so that each product and
// sum rounds as the NumPy expressions do.
__device__ __forceinline__ void camera_ray(double x, double y, double ci, double si, double d, double sx[4],
                                           double sv[4]) {
#pragma clang fp contract(off)
  const double ox = 0.0 * ci + 0.0 * si + d * si, oy = 0.0, oz = 0.0 * si + 0.0 * ci + d * ci;
  const double ix = -y * ci + 0.0 * si + d * si, iy = x, iz = y * si + 0.0 * ci + d * ci;
  const double px_ = x + y, py_ = y - x;
  const double qx = -py_ * ci + 0.0 * si + d * si, qy = px_, qz = py_ * si + 0.0 * ci + d * ci;
  const double a0 = -ix + ox, a1 = -iy + oy, a2 = -iz + oz;
  const double b0 = qx - ix, b1 = qy - iy, b2 = qz - iz;
  sx[0] = 0.0; sx[1] = ix; sx[2] = iy; sx[3] = iz;
  sv[0] = 1.0;
  sv[1] = a1 * b2 - a2 * b1;
  sv[2] = a2 * b0 - a0 * b2;
  sv[3] = a0 * b1 - a1 * b0;
}

// numpy.linspace(lo, hi, 2n+1)[1::2]
__device__ __forceinline__ double grid_coord(double lo, double step, int i) {
#pragma clang fp contract(off)
  return (double)(2 * i + 1) * step + lo;
}

// _quadratic + nullify (geodesics.py:58-85) with k^t = v0 held fixed
__device__ __forceinline__ void nullify(const double x[4], const double v[4], double a, double out[8]) {
#pragma clang fp contract(off)
  const KS<double> m = ks_metric<double>(x[1], x[2], x[3], a);
  const double l[4] = {1.0, m.l1, m.l2, m.l3};
  double g[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) g[i][j] = (i == j ? (i == 0 ? -1.0 : 1.0) : 0.0) + m.f * (l[i] * l[j]);
  const double A = v[0] * g[0][0] * v[0];
  const double b = (v[1] * g[1][0] + v[2] * g[2][0] + v[3] * g[3][0]) * v[0];
  double C = 0.0;
#pragma unroll
  for (int j = 1; j < 4; ++j) {
    const double t = v[1] * g[1][j] + v[2] * g[2][j] + v[3] * g[3][j];
    C += t * v[j];
  }
  const double bb = b * b, AC = A * C;
  const bool close = __builtin_fabs(bb - AC) <= 1e-8 + 1e-5 * __builtin_fabs(AC);
  const double dd = close ? 0.0 : bb - AC;
  const double bs = (b < 0.0) ? 0.0 : 1.0;  // heaviside(b, 1)
  const double D = -(b + bs * __builtin_sqrt(dd));
  const double x1 = D / A, x2 = C / D;
  const bool nan = (x1 != x1) || (x2 != x2);
  const double lo = nan ? __builtin_nan("") : (x1 < x2 ? x1 : x2);
  const double hi = nan ? __builtin_nan("") : (x1 > x2 ? x1 : x2);
  const double S = lo > 0.0 ? lo : (hi > 0.0 ? hi : __builtin_nan(""));
  out[0] = x[0]; out[1] = x[1]; out[2] = x[2]; out[3] = x[3];
  out[4] = v[0];
  out[5] = v[1] / S; out[6] = v[2] / S; out[7] = v[3] / S;
}

// ------------------------------------------------------------------ snapshot
// Device view of a ghost-padded meshblock snapshot.  Cells are stored as
// records of the 8 primitives in canonical order [rho, u, U1, U2, U3, B1,
// B2, B3] (32 bytes, two float4), block-major then k, j, i, so the 2 x-
// neighbours of a trilinear stencil are 64 contiguous bytes.
// Per-block geometry (kMeta doubles): lo[3] = x{1,2,3}f[0], hi[3] =
// x{1,2,3}f[-1], q0[3] = x{1,2,3}v[0] - dx, inv[3] = 1 / dx, dx = v[1] - v[0].
constexpr int kMeta = 12;

struct SnapView {
  const float4* cells;   // nmb * n2^3 * 2
  const double* meta;    // nmb * kMeta
  const int* grid;       // gz * gy * gx leaf ids (-1 = hole)
  int nmb, n2;
  int gx, gy, gz;
  double g0[3], ginv[3], dom_lo[3], dom_hi[3];
};

// Left-open, right-closed box test on a block record (lo[3], hi[3], ...),
// evaluated without short-circuit so that the six bounds load together.
__device__ __forceinline__ bool in_box(const double* r, double x, double y, double z) {
  const double l0 = r[0], l1 = r[1], l2 = r[2], h0 = r[3], h1 = r[4], h2 = r[5];
  return (l0 < x) & (x <= h0) & (l1 < y) & (y <= h1) & (l2 < z) & (z <= h2);
}

__device__ __forceinline__ bool in_block(const double* mb, double x, double y, double z) { return in_box(mb, x, y, z); }

// Slow part of the leaf lookup: the 26 neighbouring lookup cells (a point
// within rounding of a face), then a scan of every block (holes in the
// tiling), as the reference's brute-force rule.
__device__ __forceinline__ int find_block_slow(const SnapView& sv, const double* meta, const int* grid, const int c[3],
                                               double x, double y, double z) {
#pragma unroll 1
  for (int dz = -1; dz <= 1; ++dz)
#pragma unroll 1
    for (int dy = -1; dy <= 1; ++dy)
#pragma unroll 1
      for (int dx = -1; dx <= 1; ++dx) {
        const int cx = c[0] + dx, cy = c[1] + dy, cz = c[2] + dz;
        if (cx < 0 || cy < 0 || cz < 0 || cx >= sv.gx || cy >= sv.gy || cz >= sv.gz) continue;
        const int b = grid[((long)cz * sv.gy + cy) * sv.gx + cx];
        if (b >= 0 && in_block(meta + kMeta * (long)b, x, y, z)) return b;
      }
  int found = -1;
#pragma unroll 1
  for (int b = 0; b < sv.nmb; ++b)
    if (in_block(meta + kMeta * (long)b, x, y, z)) found = b;
  return found;
}

// Leaf id with the reference's rule (left-open, right-closed, last match
// wins; athenak.py:657-670), via the level-resolution lookup grid, and the
// block's record (lo, hi, q0, 1/dx) in rec.  `meta` and `grid` are the
// snapshot's tables, in global memory or copied to LDS.  The common case is
// one lookup-grid read followed by one 96-byte record read.
__device__ __forceinline__ int find_block_rec(const SnapView& sv, const double* meta, const int* grid, double x,
                                              double y, double z, double rec[kMeta]) {
  const bool inside = (sv.dom_lo[0] < x) & (x <= sv.dom_hi[0]) & (sv.dom_lo[1] < y) & (y <= sv.dom_hi[1]) &
                      (sv.dom_lo[2] < z) & (z <= sv.dom_hi[2]);
  int c[3];
  const double p[3] = {x, y, z};
  const int gn[3] = {sv.gx, sv.gy, sv.gz};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double t = inside ? (p[d] - sv.g0[d]) * sv.ginv[d] : 0.0;
    int ci = (int)__builtin_ceil(t) - 1;
    ci = ci < 0 ? 0 : (ci >= gn[d] ? gn[d] - 1 : ci);
    c[d] = ci;
  }
  int b = grid[((long)c[2] * sv.gy + c[1]) * sv.gx + c[0]];
  const double* m = meta + kMeta * (long)(b < 0 ? 0 : b);
#pragma unroll
  for (int q = 0; q < kMeta; ++q) rec[q] = m[q];
  const bool hit = inside & (b >= 0) & in_box(rec, x, y, z);
  if (!hit) {
    b = inside ? find_block_slow(sv, meta, grid, c, x, y, z) : -1;
    if (b >= 0) {
#pragma unroll
      for (int q = 0; q < kMeta; ++q) rec[q] = meta[kMeta * (long)b + q];
    }
  }
  return b;
}

__device__ __forceinline__ int find_block(const SnapView& sv, const double* meta, const int* grid, double x,
                                          double y, double z) {
  double rec[kMeta];
  return find_block_rec(sv, meta, grid, x, y, z, rec);
}

template <typename T>
__device__ __forceinline__ T lerp_ref(T a, T b, T t) {
  return a + (b - a) * t;
}

// Trilinear stencil of a point: block, first cell record, fractions
// (athenak.py:718-733: q = x - x1v[0] + dx, index floor(q/dx), fraction
// (q/dx) mod 1, in a block padded by one ghost cell).
struct Stencil {
  int mb;
  long base;  // float4 index of the (k, j, i) corner record
  double fr[3];
};

// Stencil from a known block record (mb >= 0, m its record) or none (mb < 0).
__device__ __forceinline__ Stencil stencil_in(const SnapView& sv, int mb, const double m[kMeta], double x, double y,
                                              double z) {
  Stencil st;
  st.mb = mb;
  st.base = 0;
  st.fr[0] = st.fr[1] = st.fr[2] = 0.0;
  if (st.mb < 0) return st;
  const double p[3] = {x, y, z};
  int id[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double t = (p[d] - m[6 + d]) * m[9 + d];
    const double fl = __builtin_floor(t);
    int i = (int)fl;
    i = i < 0 ? 0 : (i > sv.n2 - 2 ? sv.n2 - 2 : i);
    id[d] = i;
    st.fr[d] = t - fl;
  }
  const long n2 = sv.n2;
  st.base = 2 * (((long)st.mb * n2 + id[2]) * n2 * n2 + (long)id[1] * n2 + id[0]);
  return st;
}

__device__ __forceinline__ Stencil locate(const SnapView& sv, const double* meta, const int* grid, double x,
                                          double y, double z) {
  double m[kMeta];
  const int mb = find_block_rec(sv, meta, grid, x, y, z, m);
  return stencil_in(sv, mb, m, x, y, z);
}

// locate() with the caller's last block (cmb, its record crec; cmb < 0: none).
// Leaf blocks do not overlap (left-open, right-closed boxes), so a point
// inside the cached block lies in no other block and the table lookup is
// skipped: the same stencil as locate().
__device__ __forceinline__ Stencil locate_cached(const SnapView& sv, const double* meta, const int* grid, double x,
                                                 double y, double z, int& cmb, double crec[kMeta]) {
  if (!(cmb >= 0 && in_box(crec, x, y, z))) cmb = find_block_rec(sv, meta, grid, x, y, z, crec);
  return stencil_in(sv, cmb, crec, x, y, z);
}

// The 8 corner records of a stencil: 4 rows of 64 contiguous bytes.
struct Corners {
  float4 q[8][2];
};

__device__ __forceinline__ void fetch(const SnapView& sv, const Stencil& st, Corners& c) {
  if (st.mb < 0) return;
  const long n2 = sv.n2, sj = 2 * n2, sk = 2 * n2 * n2;
  const float4* b = sv.cells + st.base;
#pragma unroll
  for (int kk = 0; kk < 2; ++kk)
#pragma unroll
    for (int jj = 0; jj < 2; ++jj)
#pragma unroll
      for (int ii = 0; ii < 2; ++ii) {
        const float4* e = b + kk * sk + jj * sj + 2 * ii;
        c.q[kk * 4 + jj * 2 + ii][0] = e[0];
        c.q[kk * 4 + jj * 2 + ii][1] = e[1];
      }
}

// FP64: interpolate x1, then x2, then x3 (athenak.py:746-752).  FP32 (the
// mixed path): the same trilinear interpolant as a weighted sum of the 8
// corners (weights formed once, shared by the 8 primitives; 1 + 7 FMA per
// primitive instead of 7 lerps).  Whatever the stencil: callers discard the
// result of a stencil outside the domain.
template <typename T>
__device__ __forceinline__ void interp_raw(const Stencil& st, const Corners& c, T P[8]) {
  const T f0 = (T)st.fr[0], f1 = (T)st.fr[1], f2 = (T)st.fr[2];
  if constexpr (sizeof(T) == 4) {
    const T g0 = T(1) - f0, g1 = T(1) - f1, g2 = T(1) - f2;
    const T wyz[4] = {g1 * g2, f1 * g2, g1 * f2, f1 * f2};  // index x3*2 + x2
    T w[8];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      w[2 * q] = g0 * wyz[q];
      w[2 * q + 1] = f0 * wyz[q];
    }
#pragma unroll
    for (int v = 0; v < 8; ++v) {
      T acc = T(0);
#pragma unroll
      for (int c8 = 0; c8 < 8; ++c8) {
        const float4 e = c.q[c8][v >> 2];
        const int r = v & 3;
        acc += w[c8] * (r == 0 ? e.x : (r == 1 ? e.y : (r == 2 ? e.z : e.w)));
      }
      P[v] = acc;
    }
    return;
  }
#pragma unroll
  for (int v = 0; v < 8; ++v) {
    T s[8];
#pragma unroll
    for (int c8 = 0; c8 < 8; ++c8) {
      const float4 w = c.q[c8][v >> 2];
      const int r = v & 3;
      s[c8] = (T)(r == 0 ? w.x : (r == 1 ? w.y : (r == 2 ? w.z : w.w)));
    }
    // naming of the reference: d{x3}{x2}{x1}, index = x3*4 + x2*2 + x1
    const T daa = lerp_ref(s[0], s[1], f0);
    const T dab = lerp_ref(s[2], s[3], f0);
    const T dba = lerp_ref(s[4], s[5], f0);
    const T dbb = lerp_ref(s[6], s[7], f0);
    const T da = lerp_ref(daa, dab, f1);
    const T db = lerp_ref(dba, dbb, f1);
    P[v] = lerp_ref(da, db, f2);
  }
}

// Interpolated primitives; zero outside the domain.
template <typename T>
__device__ __forceinline__ void interp(const Stencil& st, const Corners& c, T P[8]) {
  if (st.mb < 0) {
#pragma unroll
    for (int k = 0; k < 8; ++k) P[k] = T(0);
    return;
  }
  interp_raw<T>(st, c, P);
}

template <typename T>
__device__ __forceinline__ void trilinear(const SnapView& sv, double x, double y, double z, T P[8]) {
  const Stencil st = locate(sv, sv.meta, sv.grid, x, y, z);
  Corners c;
  fetch(sv, st, c);
  interp<T>(st, c, P);
}

// Fluid scalars at one state (athenak.py:759-794): rho, u, cos(pitch), k.u,
// |b|.  cos(pitch) = k.b / (|k.u| |b|), NaN -> cos(fallback), clamped to
// [-1, 1]; the pitch angle itself is acos of it.  Uses g^{t mu} = eta^{t mu}
// - f l^t l^mu in closed form (alpha = 1/sqrt(1 + f), g^{ti} = f l_i).
template <typename T>
__device__ __forceinline__ void fluid_scalars(const T P[8], T x, T y, T z, T k0, T k1, T k2, T k3, T a,
                                              T fallback_cos, T out[5]) {
  using M = Math<T>;
  const KS<T> m = ks_metric<T>(x, y, z, a);
  const T f = m.f, l1 = m.l1, l2 = m.l2, l3 = m.l3;
  const T alpha = M::sqrt(M::div(T(1), T(1) + f));
  const T U1 = P[2], U2 = P[3], U3 = P[4];
  const T lU = l1 * U1 + l2 * U2 + l3 * U3;
  const T gamma = M::sqrt(T(1) + (U1 * U1 + U2 * U2 + U3 * U3 + f * lU * lU));
  const T ga = gamma * alpha;
  const T uc0 = M::div(gamma, alpha);
  const T uc1 = U1 - ga * (f * l1), uc2 = U2 - ga * (f * l2), uc3 = U3 - ga * (f * l3);
  const T Lu = uc0 + l1 * uc1 + l2 * uc2 + l3 * uc3;
  const T fLu = f * Lu;
  const T ul0 = -uc0 + fLu, ul1 = uc1 + l1 * fLu, ul2 = uc2 + l2 * fLu, ul3 = uc3 + l3 * fLu;
  const T B1 = P[5], B2 = P[6], B3 = P[7];
  const T bc0 = B1 * ul1 + B2 * ul2 + B3 * ul3;
  const T iu0 = M::div(T(1), uc0);
  const T bc1 = (B1 + uc1 * bc0) * iu0, bc2 = (B2 + uc2 * bc0) * iu0, bc3 = (B3 + uc3 * bc0) * iu0;
  const T Lb = bc0 + l1 * bc1 + l2 * bc2 + l3 * bc3;
  const T fLb = f * Lb;
  const T bl0 = -bc0 + fLb, bl1 = bc1 + l1 * fLb, bl2 = bc2 + l2 * fLb, bl3 = bc3 + l3 * fLb;
  const T kdotu = k0 * ul0 + k1 * ul1 + k2 * ul2 + k3 * ul3;
  const T kdotb = k0 * bl0 + k1 * bl1 + k2 * bl2 + k3 * bl3;
  const T bdotb = bc0 * bl0 + bc1 * bl1 + bc2 * bl2 + bc3 * bl3;
  const T bmag = M::sqrt(bdotb);
  T p = M::div(kdotb, M::abs(kdotu) * bmag);
  if (is_nan(p)) p = fallback_cos;
  if (M::abs(p) > T(1)) p = p / M::abs(p);
  out[0] = P[0];
  out[1] = P[1];
  out[2] = p;
  out[3] = kdotu;
  out[4] = bmag;
}

// ------------------------------------------------------------- emission
// transfer.py:30-86 for one sample, given sin(pitch angle).
//   nu_s = (2/9) (e B / 2 pi m_e c) Theta^2 sin(theta), X = nu / nu_s
//   j = Ne nu_s (X^1/2 + 2^(11/12) X^1/6)^2 / (2 Theta^2) e^{-X^1/3} sqrt2 pi e^2 / 3c
//   alpha = j / B_nu,  B_nu = 2 h nu^3 / (e^x - 1) / c^2,  x = h nu / m_e c^2 Theta
// X^1/6 is taken as sqrt(X^1/3).  FP32 evaluates alpha without forming
// nu^3, which overflows single precision for strongly blueshifted samples.
template <typename T>
__device__ __forceinline__ void synchrotron_sin(T Ne, T Th, T B, T sin_pitch, T nu, bool invariant, T rescale,
                                                T& jo, T& ao) {
  using M = Math<T>;
  // constant quotients folded into products; FP64 divides by nus, Th and nu
  // through three refined reciprocals (about 1 ULP, against nine IEEE
  // divisions: 90 of the kernel's 560 instructions)
  const T nuc = B * T(kEE / (2.0 * kPI * kME * kCL));
  const T nus = T(2.0 / 9.0) * nuc * (Th * Th) * sin_pitch;
  T X, ith = T(0);
  if constexpr (sizeof(T) == 8) {
    X = nu * M::rcp(nus);
    ith = M::rcp(Th);
  } else {
    X = M::div(nu, nus);
  }
  const T c3 = M::cbrt(X);
  const T var = M::exp(-c3);
  const T term = M::sqrt(X) + T(1.8877486253633870) * M::sqrt(c3);  // 2^(11/12)
  T em = sizeof(T) == 8 ? Ne * nus * (term * term) * (T(0.5) * ith * ith)
                        : M::div(Ne * nus * (term * term), T(2) * (Th * Th));
  em = em * var * T(1.4142135623730951 * kPI * kEE * kEE / (3.0 * kCL));
  if (X > T(1e12)) em = T(0);
  if (Th < T(0.3)) em = T(0);
  const T bx = sizeof(T) == 8 ? T(kHPL / (kME * kCL * kCL)) * nu * ith
                              : M::div(T(kHPL) * nu, T(kME * kCL * kCL) * Th);
  T Bden = bx * T(1.0 / 24) * (T(24) + bx * (T(12) + bx * (T(4) + bx)));
  if (!(bx < T(2e-3))) Bden = M::exp(bx) - T(1);  // rarely taken: a branch, not a select
  T ab;
  if constexpr (sizeof(T) == 8) {
    // alpha = j / B_nu with B_nu = 2 h nu^3 / (c^2 B_den)
    const T inu = M::rcp(nu);
    ab = em * Bden * T(kCL * kCL / (2.0 * kHPL)) * (inu * inu * inu);
    if (invariant) {
      const T irn = inu * M::rcp(rescale);
      em = em * (irn * irn);
      ab = ab * (nu * rescale);
    }
  } else {
    // B_nu = 2 (h nu / B_den) nu^2 / c^2: h nu / B_den stays near m_e c^2 Theta
    const T q = M::div(T(kHPL) * nu, Bden);
    const T Bnu = T(2.0 / (kCL * kCL)) * q * (nu * nu);
    ab = M::div(em, Bnu);
    if (invariant) {
      const T rn = nu * rescale;
      em = M::div(em, rn * rn);
      ab = ab * rn;
    }
  }
  if (is_nan(em)) em = T(0);
  if (is_nan(ab)) ab = T(0);
  jo = em;
  ao = ab;
}

// transfer.py:30-86 with the pitch angle itself
template <typename T>
__device__ __forceinline__ void synchrotron(T Ne, T Th, T B, T pitch, T nu, bool invariant, T rescale, T& jo,
                                            T& ao) {
  synchrotron_sin<T>(Ne, Th, B, Math<T>::sin(pitch), nu, invariant, rescale, jo, ao);
}

// electrons.py:32-50
template <typename T>
__device__ __forceinline__ T rlow_rhigh(T dens, T u, T beta, T r_low, T r_high, T ge, T gi) {
  using M = Math<T>;
  const T b2 = beta * beta;
  const T Tr = M::div(r_high * b2 + r_low, T(1) + b2);
  if constexpr (sizeof(T) == 4) {  // c^2 cancels: m_p/m_e folded into one constant
    const T te = T(kMP / kME) * u * (ge - T(1)) * (gi - T(1));
    return M::div(te, dens * ((gi - T(1)) + (ge - T(1)) * Tr));
  }
  T te = T(kCL * kCL) * (T(kMP) * u * (ge - T(1)) * (gi - T(1)));
  te = M::div(te, dens * ((gi - T(1)) + (ge - T(1)) * Tr));
  return te / T(kME * kCL * kCL);
}

// Per-sample image parameters (images.py:86-118, grmhd.py:40-55)
struct ImageParams {
  double fluid_gamma, r_high, Ne_unit, B_unit, L_unit, nu_obs, fallback_cos, sigma_cut;
};

// (j, alpha) from interpolated primitives at one accepted state
// (images.py:86-118).  sin(acos(p)) is taken as sqrt(1 - p^2).
template <typename T>
__device__ __forceinline__ void emission_from_prims(const ImageParams& ip, const T P[8], const double s[8], double a,
                                                    T& j, T& al) {
  using M = Math<T>;
  T sc[5];
  fluid_scalars<T>(P, (T)s[1], (T)s[2], (T)s[3], (T)s[4], (T)s[5], (T)s[6], (T)s[7], (T)a, (T)ip.fallback_cos, sc);
  const T dens = sc[0], u = sc[1], cp = sc[2], kdotu = sc[3], b = sc[4];
  const T bsq = b * b;
  const T beta = M::div(u * (T(ip.fluid_gamma) - T(1)), bsq) / T(0.5);
  const T sigma = M::div(bsq, dens);
  const T Th = rlow_rhigh<T>(dens, u, beta, T(1), T(ip.r_high), T(4.0 / 3), T(5.0 / 3));
  const T Ne = T(ip.Ne_unit) * dens, Bg = T(ip.B_unit) * b, nu = -kdotu * T(ip.nu_obs);
  const T one_m = T(1) - cp * cp;
  const T sp = M::sqrt(one_m > T(0) ? one_m : T(0));
  synchrotron_sin<T>(Ne, Th, Bg, sp, nu, true, T(1.0 / ip.nu_obs), j, al);
  if (sigma > T(ip.sigma_cut)) {
    j = T(0);
    al = T(0);
  }
}

}  // namespace mk
