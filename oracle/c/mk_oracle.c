/*
 * mk_oracle.c -- plain C FP64 restatement of Mahakala's per-ray hot path.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py (as the CPU baseline that is timed).  The
 * product library never links it.
 *
 * It restates (file:line in the reference's mahakala/ package):
 *   camera grid / pixels      geodesics.py:107-134, 137-216, 204-216
 *   nullify                   geodesics.py:58-85, 219-230
 *   metric and its gradient   geodesics.py:88-104, 305 (closed-form gradient
 *                             of the same expression; inverse in closed form
 *                             g^{mu nu} = eta - f l^mu l^nu instead of LU)
 *   rhs / RK4                 geodesics.py:294-336
 *   step rule, freeze, trim   geodesics.py:233-291, 350-351
 *   photon selection          geodesics.py:354-378 (argmax(dt)-1, -1 wraps)
 *   bisection                 geodesics.py:405-435
 *   sampler                   grmhd/athenak.py:639-812
 *   electrons / units         electrons.py:32-50, grmhd/grmhd.py:40-55
 *   synchrotron j, alpha      transfer.py:30-86
 *   transfer                  transfer.py:89-119 (accumulated forward along
 *                             the ray: I = sum_i b_i prod_{k<i} a_k, the same
 *                             linear recurrence the backward scan evaluates)
 *   image glue                images.py:81-120
 *
 * It is pinned against oracle/mahakala_np.py (the line-by-line NumPy
 * restatement) on small inputs, and against closed forms in tests/.
 * Build: oracle/Makefile (gcc -O3 -fopenmp -ffp-contract=off).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* constants.py:23-31 (+ grmhd.py:26-32, electrons.py:23-29) */
#define EE 4.8032e-10
#define CL 2.99792458e10
#define ME 9.1094e-28
#define MP 1.6726e-24
#define HPL 6.6261e-27
#define GNEWT 6.6743e-8
#define PI_D 3.141592653589793

/* ------------------------------------------------------------------ metric */
typedef struct {
    double f, l[4];     /* covariant l_mu, l_0 = 1 */
    double df[4];       /* d_sigma f, sigma = 0..3 (d_t = 0) */
    double dl[4][4];    /* dl[mu][sigma] = d_sigma l_mu */
} ks_t;

static void ks_eval(const double* x, double a, ks_t* m, int want_grad) {
    double X = x[1], Y = x[2], Z = x[3];
    double aa = a * a, zz = Z * Z;
    double kk = 0.5 * (X * X + Y * Y + zz - aa);
    double Q = sqrt(kk * kk + aa * zz);
    double rr = Q + kk;
    double r = sqrt(rr);
    double den = rr * rr + aa * zz;
    double f = (2.0 * rr * r) / den;
    double s = 1.0 / (rr + aa);
    m->f = f;
    m->l[0] = 1.0;
    m->l[1] = (r * X + a * Y) * s;
    m->l[2] = (r * Y - a * X) * s;
    m->l[3] = Z / r;
    if (!want_grad) return;
    /* d rr / d x_i = x rr/Q, y rr/Q, z (rr + a^2)/Q ; d r = d rr / (2 r) */
    double drr[4] = {0.0, X * rr / Q, Y * rr / Q, Z * (rr + aa) / Q};
    double dr[4];
    for (int i = 0; i < 4; ++i) dr[i] = drr[i] / (2.0 * r);
    /* f = 2 r^3 / (r^4 + a^2 z^2) */
    double r2 = rr, r3 = rr * r;
    for (int i = 0; i < 4; ++i) {
        double num = 6.0 * r2 * dr[i] * den - 2.0 * r3 * (2.0 * rr * drr[i] + (i == 3 ? 2.0 * aa * Z : 0.0));
        m->df[i] = num / (den * den);
    }
    m->df[0] = 0.0;
    for (int i = 0; i < 4; ++i) {
        m->dl[0][i] = 0.0;
        double dx = (i == 1), dy = (i == 2), dz = (i == 3);
        m->dl[1][i] = s * (dr[i] * X + r * dx + a * dy - m->l[1] * drr[i]);
        m->dl[2][i] = s * (dr[i] * Y + r * dy - a * dx - m->l[2] * drr[i]);
        m->dl[3][i] = (dz - m->l[3] * dr[i]) / r;
    }
    for (int mu = 0; mu < 4; ++mu) m->dl[mu][0] = 0.0;
}

void mko_metric(const double* x, double a, double* g) {
    ks_t m;
    ks_eval(x, a, &m, 0);
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) g[4 * i + j] = (i == j ? (i == 0 ? -1.0 : 1.0) : 0.0) + m.f * (m.l[i] * m.l[j]);
}

/* a^alpha = -g^{ab} (d_s g_{bn} v^s v^n - 1/2 d_b g_{mn} v^m v^n)  (geodesics.py:294-309) */
static void rhs8(const double* st, double a, double* out) {
    ks_t m;
    ks_eval(st, a, &m, 1);
    const double* v = st + 4;
    double L = 0, Fv = 0, W[4] = {0}, D[4] = {0};
    for (int mu = 0; mu < 4; ++mu) L += m.l[mu] * v[mu];
    for (int s = 0; s < 4; ++s) Fv += m.df[s] * v[s];
    for (int mu = 0; mu < 4; ++mu)
        for (int s = 0; s < 4; ++s) {
            W[mu] += m.dl[mu][s] * v[s];
            D[s] += m.dl[mu][s] * v[mu];
        }
    double Wv = 0;
    for (int mu = 0; mu < 4; ++mu) Wv += W[mu] * v[mu];
    double c[4];
    for (int b = 0; b < 4; ++b) {
        double A = m.l[b] * (L * Fv + m.f * Wv) + m.f * L * W[b];
        double B = 0.5 * m.df[b] * L * L + m.f * L * D[b];
        c[b] = B - A;
    }
    double lc = -c[0] + m.l[1] * c[1] + m.l[2] * c[2] + m.l[3] * c[3];
    for (int i = 0; i < 4; ++i) out[i] = v[i];
    out[4] = -c[0] + m.f * lc;
    for (int i = 1; i < 4; ++i) out[4 + i] = c[i] - m.f * m.l[i] * lc;
}

void mko_rhs(const double* s, double a, double* out, long n) {
    for (long i = 0; i < n; ++i) rhs8(s + 8 * i, a, out + 8 * i);
}

static void rk4(const double* s, double dt, double a, double* out) {
    double k1[8], k2[8], k3[8], k4[8], t[8];
    rhs8(s, a, k1);
    for (int i = 0; i < 8; ++i) { k1[i] *= dt; t[i] = s[i] + 0.5 * k1[i]; }
    rhs8(t, a, k2);
    for (int i = 0; i < 8; ++i) { k2[i] *= dt; t[i] = s[i] + 0.5 * k2[i]; }
    rhs8(t, a, k3);
    for (int i = 0; i < 8; ++i) { k3[i] *= dt; t[i] = s[i] + k3[i]; }
    rhs8(t, a, k4);
    for (int i = 0; i < 8; ++i) { k4[i] *= dt; out[i] = s[i] + 1.0 / 6.0 * (k1[i] + 2.0 * k2[i] + 2.0 * k3[i] + k4[i]); }
}

void mko_rk4(const double* s, const double* dt, double a, double* out, long n) {
    for (long i = 0; i < n; ++i) rk4(s + 8 * i, dt[i], a, out + 8 * i);
}

/* geodesics.py:284-291 */
static double radius_cal(const double* x, double a) {
    double R = sqrt(x[1] * x[1] + x[2] * x[2] + x[3] * x[3]);
    double R2 = R * R, aa = a * a;
    return sqrt((R2 - aa + sqrt((R2 - aa) * (R2 - aa) + 4.0 * aa * (x[3] * x[3]))) / 2.0);
}

/* geodesics.py:249-252 */
static double step_rule(const double* x, double a, double rEH, double div, double tol) {
    double dt = -(radius_cal(x, a) - rEH) / div;
    if (isnan(dt) || fabs(dt) * div < tol || fabs(dt) * div > 1500.0) return 0.0;
    return dt;
}

/* Relative distance of a step-rule magnitude m = |r - r_EH| from the nearest
 * freeze threshold (tol or 1500), infinite for NaN: a ray whose stop
 * decision differs between two FP64 evaluations of the same trajectory must
 * have met a candidate within rounding of a threshold. */
static double threshold_margin(const double* x, double a, double rEH, double tol) {
    double m = fabs(radius_cal(x, a) - rEH);
    if (isnan(m)) return INFINITY;
    double d1 = fabs(m - tol) / tol, d2 = fabs(m - 1500.0) / 1500.0;
    return d1 < d2 ? d1 : d2;
}

/* ------------------------------------------------------------------ camera */
static void image_to_bh(double x, double y, double z, double ci, double si, double d, double* o) {
    o[0] = -y * ci + z * si + d * si;
    o[1] = x;
    o[2] = y * si + z * ci + d * ci;
}

static void camera_one(double x, double y, double inc, double d, double* sx, double* sv) {
    double i = inc * PI_D / 180.0;
    double ci = cos(i), si = sin(i);
    double org[3], ini[3], per[3], v1[3], v2[3];
    image_to_bh(0.0, 0.0, 0.0, ci, si, d, org);
    image_to_bh(x, y, 0.0, ci, si, d, ini);
    image_to_bh(x + y, y - x, 0.0, ci, si, d, per);
    for (int k = 0; k < 3; ++k) { v1[k] = -ini[k] + org[k]; v2[k] = per[k] - ini[k]; }
    sx[0] = 0.0; sx[1] = ini[0]; sx[2] = ini[1]; sx[3] = ini[2];
    sv[0] = 1.0;
    sv[1] = v1[1] * v2[2] - v1[2] * v2[1];
    sv[2] = v1[2] * v2[0] - v1[0] * v2[2];
    sv[3] = v1[0] * v2[1] - v1[1] * v2[0];
}

static int isclose(double a, double b) { return fabs(a - b) <= 1e-8 + 1e-5 * fabs(b); }

/* geodesics.py:58-85 */
static void nullify(const double* x, const double* v, double a, double* out) {
    double g[16];
    mko_metric(x, a, g);
    double A = v[0] * g[0] * v[0];
    double b = (v[1] * g[4] + v[2] * g[8] + v[3] * g[12]) * v[0];
    double C = 0;
    for (int j = 1; j < 4; ++j) {
        double t = 0;
        for (int i = 1; i < 4; ++i) t += v[i] * g[4 * i + j];
        C += t * v[j];
    }
    double bb = b * b, AC = A * C;
    double dd = isclose(bb, AC) ? 0.0 : bb - AC;
    double bs = (b < 0) ? 0.0 : 1.0; /* heaviside(b, 1) */
    double D = -(b + bs * sqrt(dd));
    double x1 = D / A, x2 = C / D;
    double lo = (isnan(x1) || isnan(x2)) ? NAN : (x1 < x2 ? x1 : x2);
    double hi = (isnan(x1) || isnan(x2)) ? NAN : (x1 > x2 ? x1 : x2);
    double S = lo > 0 ? lo : (hi > 0 ? hi : NAN);
    for (int i = 0; i < 4; ++i) out[i] = x[i];
    out[4] = v[0];
    for (int i = 1; i < 4; ++i) out[4 + i] = v[i] / S;
}

/* grid pixel centres: numpy linspace(lo, hi, 2n+1)[1::2], pixel p = i n + j */
static double grid_coord(double lo, double hi, int n, int i) {
    double step = (hi - lo) / (double)(2 * n);
    return (double)(2 * i + 1) * step + lo;
}

void mko_camera_grid(double a, double inc, double dist, double lo, double hi, int n, double* out) {
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            double sx[4], sv[4];
            camera_one(grid_coord(lo, hi, n, i), grid_coord(lo, hi, n, j), inc, dist, sx, sv);
            nullify(sx, sv, a, out + 8 * ((long)i * n + j));
        }
}

void mko_camera_pixels(double a, double inc, double dist, const double* radius, const double* angle, long n, double* out) {
    for (long p = 0; p < n; ++p) {
        double sx[4], sv[4];
        camera_one(cos(angle[p]) * radius[p], sin(angle[p]) * radius[p], inc, dist, sx, sv);
        nullify(sx, sv, a, out + 8 * p);
    }
}

/* ------------------------------------------------------------- integrator */
typedef struct {
    int K;            /* first frozen row, N if never */
    double fin[8];    /* state of the ray's last emitted row (row min(K, N-1)) */
    double sel_r;     /* select_photons_integrator value for this ray */
    double m_rej;     /* threshold_margin of the candidate that froze the ray */
    double m_acc;     /* threshold_margin of the last accepted state */
} ray_out_t;

/* Row n of the reference's history is (s_n, dt_n); a ray frozen at row K
 * emits (s_K, 0) on every row >= K.  The value select_photons_integrator
 * reads is r at row argmax(dt) - 1, where -1 wraps to the last row of the
 * trimmed history, which for this ray holds its last emitted state. */
static void integrate_one(const double* s0, int N, double div, double tol, double a, ray_out_t* o) {
    double rEH = 1.0 + sqrt(1.0 - a * a);
    double s[8], nw[8];
    memcpy(s, s0, sizeof s);
    double dt = step_rule(s, a, rEH, div, tol);
    double best = -INFINITY, r_prev = NAN, r_at_best_prev = NAN;
    int besti = -1, K = N;
    o->m_rej = INFINITY;
    for (int n = 0; n < N; ++n) {
        double r_here = radius_cal(s, a);
        int frozen = (dt == 0.0);
        double dtn = 0.0;
        if (!frozen) {
            rk4(s, dt, a, nw);
            dtn = step_rule(nw, a, rEH, div, tol);
            frozen = (dtn == 0.0);
            if (frozen) o->m_rej = threshold_margin(nw, a, rEH, tol);
        }
        double row_dt = frozen ? 0.0 : dt;
        if (row_dt > best) { best = row_dt; besti = n; r_at_best_prev = r_prev; }
        if (frozen) { K = n; break; }
        r_prev = r_here;
        if (n == N - 1) break; /* row N-1 is the last row emitted */
        memcpy(s, nw, sizeof s);
        dt = dtn;
    }
    o->K = K;
    memcpy(o->fin, s, sizeof s);
    o->m_acc = K >= N ? INFINITY : threshold_margin(s, a, rEH, tol);
    if (K >= N) o->m_rej = INFINITY;
    o->sel_r = (besti >= 1) ? r_at_best_prev : radius_cal(s, a);
}

/* final-state integration: K per ray and the last emitted state; margin
 * (NULL: not written) 2 per ray as in mko_render */
void mko_integrate_final(const double* s0, long npx, int N, double div, double tol, double a,
                         int* K, double* fin, double* sel_r, double* margin, int nthreads) {
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
    for (long p = 0; p < npx; ++p) {
        ray_out_t o;
        integrate_one(s0 + 8 * p, N, div, tol, a, &o);
        if (K) K[p] = o.K;
        if (fin) memcpy(fin + 8 * p, o.fin, sizeof o.fin);
        if (sel_r) sel_r[p] = o.sel_r;
        if (margin) { margin[2 * p] = o.m_rej; margin[2 * p + 1] = o.m_acc; }
    }
}

/* materialised history rows 0..T-1 (T given), layout (T, npx, 8) and (T, npx) */
void mko_integrate_history(const double* s0, long npx, int N, double div, double tol, double a,
                           long T, double* S, double* D, int nthreads) {
    double rEH = 1.0 + sqrt(1.0 - a * a);
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
    for (long p = 0; p < npx; ++p) {
        double s[8], nw[8];
        memcpy(s, s0 + 8 * p, sizeof s);
        double dt = step_rule(s, a, rEH, div, tol);
        int frozen = 0;
        for (long n = 0; n < T; ++n) {
            double edt = 0.0;
            if (!frozen) {
                if (dt == 0.0) frozen = 1;
                else {
                    rk4(s, dt, a, nw);
                    double dtn = step_rule(nw, a, rEH, div, tol);
                    if (dtn == 0.0) frozen = 1;
                    else edt = dt;
                }
            }
            memcpy(S + 8 * (n * npx + p), s, sizeof s);
            D[n * npx + p] = edt;
            if (!frozen) { memcpy(s, nw, sizeof s); dt = step_rule(s, a, rEH, div, tol); }
        }
        (void)N;
    }
}

/* geodesics.py:405-435: every angle shares one bracket width, so a per-ray
 * loop with the iteration count fixed up front is the same computation. */
int mko_bisect_iterations(double error_allowed, int max_it) {
    double inner = 0.5, outer = 10.0;
    int it = 0;
    while (outer - inner > error_allowed && it < max_it) { inner = (outer - inner) / 2 + inner; ++it; }
    return it;
}

void mko_shadow_bisect(double a, double inc, const double* angles, long n, int max_steps,
                       double error_allowed, int max_it, double dist, double* radii, int nthreads) {
    int iters = mko_bisect_iterations(error_allowed, max_it);
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
    for (long p = 0; p < n; ++p) {
        double inner = 0.5, outer = 10.0;
        for (int it = 0; it < iters; ++it) {
            double mid = (outer - inner) / 2 + inner;
            double sx[4], sv[4], s0[8];
            camera_one(cos(angles[p]) * mid, sin(angles[p]) * mid, inc, dist, sx, sv);
            nullify(sx, sv, a, s0);
            ray_out_t o;
            integrate_one(s0, max_steps, 40.0, 1e-2, a, &o);
            if (o.sel_r < 100.0) inner = mid; else outer = mid;
        }
        radii[p] = inner;
    }
}

/* ---------------------------------------------------------------- sampler */
/* Snapshot: blocks (nmb, 8, n+2, n+2, n+2) float32 in canonical prim order
 * [rho, u, U1, U2, U3, B1, B2, B3]; per block x{1,2,3}f bounds, x{1,2,3}v[0]
 * and dx.  Lookup: uniform grid of leaf ids at the finest block size,
 * verified against the block bounds (left-open, right-closed) with a brute
 * force fallback, so the result equals the reference's last-match scan. */
typedef struct {
    const float* data;
    long nmb;
    int n;
    const double* lo;   /* (nmb, 3) x1f[0], x2f[0], x3f[0] */
    const double* hi;   /* (nmb, 3) x1f[-1], ... */
    const double* v0;   /* (nmb, 3) x1v[0], ... */
    const double* dx;   /* (nmb, 3) x1v[1]-x1v[0], ... */
    const int* grid;    /* (G, G, G) leaf id, -1 outside */
    int G;
    double g0[3], gh;   /* grid origin and cell width */
} snap_t;

static int in_block(const snap_t* sn, long b, const double* x) {
    for (int d = 0; d < 3; ++d)
        if (!(sn->lo[3 * b + d] < x[1 + d] && x[1 + d] <= sn->hi[3 * b + d])) return 0;
    return 1;
}

static long find_block(const snap_t* sn, const double* x) {
    if (sn->grid) {
        int c[3];
        int ok = 1;
        for (int d = 0; d < 3; ++d) {
            double t = (x[1 + d] - sn->g0[d]) / sn->gh;
            if (!(t > -1.0 && t < sn->G + 1.0)) { ok = 0; break; }
            int ci = (int)ceil(t) - 1;
            if (ci < 0) ci = 0;
            if (ci >= sn->G) ci = sn->G - 1;
            c[d] = ci;
        }
        if (ok) {
            long b = sn->grid[((long)c[2] * sn->G + c[1]) * sn->G + c[0]];
            if (b >= 0 && in_block(sn, b, x)) return b;
        } else {
            return -1;
        }
    }
    long found = -1;
    for (long b = 0; b < sn->nmb; ++b)
        if (in_block(sn, b, x)) found = b;
    return found;
}

static void prims_at(const snap_t* sn, const double* x, double* P, long* mbo) {
    long mb = find_block(sn, x);
    if (mbo) *mbo = mb;
    if (mb < 0) { for (int k = 0; k < 8; ++k) P[k] = 0.0; return; }
    int n2 = sn->n + 2;
    long idx[3]; double fr[3];
    for (int d = 0; d < 3; ++d) {
        double q = x[1 + d] - sn->v0[3 * mb + d] + sn->dx[3 * mb + d];
        double t = q / sn->dx[3 * mb + d];
        double fl = floor(t);
        idx[d] = (long)fl;
        fr[d] = t - fl;
        if (idx[d] < 0) idx[d] = 0;
        if (idx[d] > n2 - 2) idx[d] = n2 - 2;
    }
    const float* base = sn->data + (size_t)mb * 8 * n2 * n2 * n2;
    for (int k = 0; k < 8; ++k) {
        const float* c = base + (size_t)k * n2 * n2 * n2;
#define AT(a3, a2, a1) ((double)c[((idx[2] + (a3)) * n2 + (idx[1] + (a2))) * n2 + idx[0] + (a1)])
        double daa = AT(0, 0, 0) + (AT(0, 0, 1) - AT(0, 0, 0)) * fr[0];
        double dab = AT(0, 1, 0) + (AT(0, 1, 1) - AT(0, 1, 0)) * fr[0];
        double dba = AT(1, 0, 0) + (AT(1, 0, 1) - AT(1, 0, 0)) * fr[0];
        double dbb = AT(1, 1, 0) + (AT(1, 1, 1) - AT(1, 1, 0)) * fr[0];
#undef AT
        double da = daa + (dab - daa) * fr[1];
        double db = dba + (dbb - dba) * fr[1];
        P[k] = da + (db - da) * fr[2];
    }
}

/* athenak.py:759-794: scalars [rho, u, pitch, k.u, |b|] */
static void scalars_at(const snap_t* sn, const double* st, double a, double fallback, double* out) {
    double P[8];
    prims_at(sn, st, P, NULL);
    ks_t m;
    ks_eval(st, a, &m, 0);
    double f = m.f; const double* l = m.l;
    double alpha = sqrt(1.0 / (1.0 + f));
    double lU = l[1] * P[2] + l[2] * P[3] + l[3] * P[4];
    double UU = P[2] * P[2] + P[3] * P[3] + P[4] * P[4] + f * lU * lU;
    double gamma = sqrt(1.0 + UU);
    double uc[4], ul[4], bc[4], bl[4];
    uc[0] = gamma / alpha;
    for (int i = 1; i < 4; ++i) uc[i] = P[1 + i] - gamma * alpha * (f * l[i]);
    double Lu = uc[0] + l[1] * uc[1] + l[2] * uc[2] + l[3] * uc[3];
    ul[0] = -uc[0] + f * Lu;
    for (int i = 1; i < 4; ++i) ul[i] = uc[i] + f * l[i] * Lu;
    bc[0] = P[5] * ul[1] + P[6] * ul[2] + P[7] * ul[3];
    for (int i = 1; i < 4; ++i) bc[i] = (P[4 + i] + uc[i] * bc[0]) / uc[0];
    double Lb = bc[0] + l[1] * bc[1] + l[2] * bc[2] + l[3] * bc[3];
    bl[0] = -bc[0] + f * Lb;
    for (int i = 1; i < 4; ++i) bl[i] = bc[i] + f * l[i] * Lb;
    const double* k = st + 4;
    double kdotu = 0, kdotb = 0, bdotb = 0;
    for (int i = 0; i < 4; ++i) { kdotu += k[i] * ul[i]; kdotb += k[i] * bl[i]; bdotb += bc[i] * bl[i]; }
    double p = kdotb / (fabs(kdotu) * sqrt(bdotb));
    if (isnan(p)) p = cos(fallback);
    if (fabs(p) > 1.0) p = p / fabs(p);
    out[0] = P[0]; out[1] = P[1]; out[2] = acos(p); out[3] = kdotu; out[4] = sqrt(bdotb);
}

void mko_sample_scalars(const float* data, long nmb, int n, const double* lo, const double* hi,
                        const double* v0, const double* dx, const int* grid, int G, const double* g0, double gh,
                        const double* S, long ns, double a, double fallback, double* out5, long* mb_out, int nthreads) {
    snap_t sn = {data, nmb, n, lo, hi, v0, dx, grid, G, {g0 ? g0[0] : 0, g0 ? g0[1] : 0, g0 ? g0[2] : 0}, gh};
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (long i = 0; i < ns; ++i) {
        double o[5];
        scalars_at(&sn, S + 8 * i, a, fallback, o);
        for (int k = 0; k < 5; ++k) out5[5 * i + k] = o[k];
        if (mb_out) { double P[8]; prims_at(&sn, S + 8 * i, P, &mb_out[i]); }
    }
}

/* ----------------------------------------------------------- emission/RT */
/* transfer.py:30-86 for one sample */
static void synch(double Ne, double Th, double B, double pitch, double nu, int invariant, double rescale,
                  double* jo, double* ao) {
    double nuc = EE * B / (2.0 * PI_D * ME * CL);
    double nus = (2.0 / 9.0) * nuc * (Th * Th) * sin(pitch);
    double X = nu / nus;
    double var = exp(-pow(X, 1.0 / 3));
    double term = sqrt(X) + pow(2.0, 11.0 / 12) * pow(X, 1.0 / 6);
    double em = Ne * nus * (term * term) / (2.0 * (Th * Th));
    em = em * var * sqrt(2.0) * PI_D * (EE * EE) / (3.0 * CL);
    if (X > 1e12) em = 0.0;
    if (Th < 0.3) em = 0.0;
    double bx = HPL * nu / (ME * CL * CL * Th);
    double series = bx / 24.0 * (24.0 + bx * (12.0 + bx * (4.0 + bx)));
    double Bden = (bx < 2e-3) ? series : (exp(bx) - 1);
    double Bnu = (2.0 * HPL * pow(nu, 3.0) / Bden) / (CL * CL);
    double ab = em / Bnu;
    if (invariant) {
        double rn = nu * rescale;
        em = em / (rn * rn);
        ab = ab * rn;
    }
    if (isnan(em)) em = 0.0;
    if (isnan(ab)) ab = 0.0;
    *jo = em; *ao = ab;
}

void mko_synchrotron(const double* Ne, const double* Th, const double* B, const double* pitch, const double* nu,
                     long n, int invariant, double rescale, double* j, double* al) {
    for (long i = 0; i < n; ++i) synch(Ne[i], Th[i], B[i], pitch[i], nu[i], invariant, rescale, j + i, al + i);
}

typedef struct {
    double fluid_gamma, r_high, Ne_unit, B_unit, L_unit, nu_obs, fallback;
} img_par_t;

/* images.py:86-118 for one sample -> (j, alpha) */
static void emission(const snap_t* sn, const double* st, double a, const img_par_t* ip, double* j, double* al) {
    double sc[5];
    scalars_at(sn, st, a, ip->fallback, sc);
    double dens = sc[0], u = sc[1], pitch = sc[2], kdotu = sc[3], b = sc[4];
    double bsq = b * b;
    double beta = u * (ip->fluid_gamma - 1.0) / bsq / 0.5;
    double sigma = bsq / dens;
    /* electrons.py:32-50 with r_low=1, gamma_e=4/3, gamma_i=5/3 */
    double ge = 4.0 / 3, gi = 5.0 / 3;
    double Tr = (ip->r_high * (beta * beta) + 1.0) / (1.0 + beta * beta);
    double te = (CL * CL) * (MP * u * (ge - 1.0) * (gi - 1.0));
    te = te / (dens * ((gi - 1.0) + (ge - 1.0) * Tr));
    double Th = te / (ME * CL * CL);
    double Ne = ip->Ne_unit * dens, Bg = ip->B_unit * b, nu = -kdotu * ip->nu_obs;
    synch(Ne, Th, Bg, pitch, nu, 1, 1.0 / ip->nu_obs, j, al);
    if (sigma > 100.0) { *j = 0.0; *al = 0.0; }
}

/* Whole image: camera -> geodesic -> sampling -> emission -> transfer.
 * pixels p0 + q * stride, q < count, of the res x res grid; I (count),
 * accepted steps per ray.
 * force_steps (NULL: the reference's rule) runs exactly force_steps[q]
 * accepted steps where the ray stays finite: a candidate the freeze rule
 * would reject is accepted with the unclamped step -(r - r_EH)/div, and the
 * ray stops when the count is reached.  Used to compare an image whose stop
 * decisions differ by rounding at a threshold with the same step counts.
 * margin (NULL: not written, 2 per ray): threshold_margin of the rejected
 * candidate that froze the ray, and of the ray's last accepted state
 * (infinite where the ray ran into the step cap). */
void mko_render(const float* data, long nmb, int n, const double* lo, const double* hi, const double* v0,
                const double* dx, const int* grid, int G, const double* g0, double gh,
                double a, double inc, double dist, double fov, int res, int max_nsteps,
                double fluid_gamma, double r_high, double mass_scale, double M_bh, double nu_obs,
                long p0, long count, long stride, double* I_out, int* steps_out, const int* force_steps,
                double* margin, int nthreads) {
    snap_t sn = {data, nmb, n, lo, hi, v0, dx, grid, G, {g0[0], g0[1], g0[2]}, gh};
    double L_unit = GNEWT * M_bh / (CL * CL);
    double dens_unit = mass_scale / (L_unit * L_unit * L_unit);
    img_par_t ip = {fluid_gamma, r_high, dens_unit / (MP + ME), CL * sqrt(4.0 * PI_D * dens_unit), L_unit, nu_obs,
                    PI_D / 3.0};
    double rEH = 1.0 + sqrt(1.0 - a * a);
    const double div = 40.0, tol = 1e-4;
#pragma omp parallel for schedule(dynamic, 8) num_threads(nthreads)
    for (long q = 0; q < count; ++q) {
        long p = p0 + q * stride;
        int i = (int)(p / res), jj = (int)(p % res);
        double sx[4], sv[4], s[8], nw[8];
        camera_one(grid_coord(-fov / 2.0, fov / 2.0, res, i), grid_coord(-fov / 2.0, fov / 2.0, res, jj), inc, dist,
                   sx, sv);
        nullify(sx, sv, a, s);
        double dt = step_rule(s, a, rEH, div, tol);
        const int force = force_steps ? force_steps[q] : -1;
        if (force > 0 && dt == 0.0) dt = -(radius_cal(s, a) - rEH) / div;
        double I = 0.0, P = 1.0, m_rej = INFINITY;
        int steps = 0;
        for (int k = 0; k < max_nsteps; ++k) {
            if (dt == 0.0 || dt != dt) break;
            if (force >= 0 && steps >= force) break;
            rk4(s, dt, a, nw);
            double dtn = step_rule(nw, a, rEH, div, tol);
            if (dtn == 0.0) {
                m_rej = threshold_margin(nw, a, rEH, tol);
                if (force >= 0 && steps + 1 < force) dtn = -(radius_cal(nw, a) - rEH) / div;
                if (force < 0 || steps + 1 > force || dtn != dtn) break;
            }
            ++steps;
            if (k + 1 <= max_nsteps - 1) {
                double jv, av;
                emission(&sn, nw, a, &ip, &jv, &av);
                double c = -dt * L_unit;
                I += P * (c * jv);
                P *= (1.0 - c * av);
            }
            memcpy(s, nw, sizeof s);
            dt = dtn;
        }
        I_out[q] = I;
        if (steps_out) steps_out[q] = steps;
        if (margin) {
            int capped = steps >= max_nsteps;
            margin[2 * q] = capped ? INFINITY : m_rej;
            margin[2 * q + 1] = capped ? INFINITY : threshold_margin(s, a, rEH, tol);
        }
    }
}

int mko_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
