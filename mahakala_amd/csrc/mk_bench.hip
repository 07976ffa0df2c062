// mk_bench.hip -- one-wave timing of the per-step device functions of the
// fused image kernel (a measurement aid; no reference counterpart).
//
// One wave runs alone on one SIMD and repeats one piece of the per-ray loop
// `steps` times; the shader clock (s_memtime) and the 100 MHz real-time
// counter around the loop give cycles and nanoseconds per repetition.  The
// pieces are the stages of mk_render's per-ray work in each precision:
//   0 / 1  RK4 step, FP64 / mixed (rk4_step<double> / rk4_step_mixed)
//   2 / 3  geodesic step: RK4 + radius + step rule, FP64 / mixed (the
//          geodesic wave's arithmetic per round, geodesics.py:249-261, 317-336)
//   4 / 5  trilinear interpolation of the 8 primitives from 8 corner records
//          read from LDS, FP64 / FP32 (athenak.py:737-752)
//   6 / 7  emission of one sample: fluid scalars, R(beta), j and alpha,
//          FP64 / FP32 (athenak.py:759-794, images.py:86-118,
//          transfer.py:30-86)
// A wave alone issues one VALU instruction per 4 cycles at best (64 lanes on
// 16-lane SIMD units), so the cycles of a stage are the latency that stage
// adds to the chain of a ray, and their sum over the stages is what one
// accepted step costs a SIMD.
#include <hip/hip_runtime.h>

#include <cmath>

#include "mk_device.h"
#include "mk_internal.h"

namespace mk {

// Forces v to be computed before what follows in program order.
__device__ __forceinline__ void pin(double v) {
  const int x = __builtin_amdgcn_readfirstlane((int)(v > 1e300));
  asm volatile("s_nop 0" ::"s"(x));
}

__global__ void __launch_bounds__(64) k_step_cycles(int variant, int steps, double a, double* out,
                                                    unsigned long long* t) {
  const int lane = threadIdx.x;
  // a tile of 8x8 camera rays across the photon ring of a 256^2, +-10 M image
  // at inclination 30 deg
  const double x_img = 4.8 + 0.078 * (lane >> 3), y_img = -0.3 + 0.078 * (lane & 7);
  const double inc = 30.0 * 3.141592653589793 / 180.0;
  double sx[4], sv4[4], s[8];
  camera_ray(x_img, y_img, __builtin_cos(inc), __builtin_sin(inc), 1000.0, sx, sv4);
  nullify(sx, sv4, a, s);
  // move the ray in to a few M, so that the timed steps are those near the hole
  const double rEH = 1.0 + __builtin_sqrt(1.0 - a * a);
  for (int k = 0; k < 200; ++k) {
    double nw[8];
    rk4_step<double>(s, -(radius_cal(s[1], s[2], s[3], a) - rEH) / 40.0, a, nw);
    for (int q = 0; q < 8; ++q) s[q] = nw[q];
  }
  double acc = 0.0;
  unsigned long long c0 = 0, c1 = 0, w0 = 0, w1 = 0;
  if (variant <= 3) {
    double dt = -0.02;
    pin(s[1]);
    c0 = __builtin_readcyclecounter();
    w0 = __builtin_amdgcn_s_memrealtime();
    for (int k = 0; k < steps; ++k) {
      double nw[8];
      if (variant == 0 || variant == 2) rk4_step<double>(s, dt, a, nw);
      else rk4_step_mixed(s, dt, (float)a, nw);
      if (variant >= 2) {
        double rn, dtn;
        if (variant == 2) {
          rn = radius_cal(nw[1], nw[2], nw[3], a);
          dtn = step_rule(rn, rEH, 40.0, 1e-4);
        } else {
          rn = radius_f32(nw[1], nw[2], nw[3], a);
          dtn = step_rule_fast(rn, rEH, 1.0 / 40.0, 40.0, 1e-4);
        }
        // keep stepping (the timed loop has no exit)
        dt = dtn != 0.0 ? dtn : -0.02;
      }
#pragma unroll
      for (int q = 0; q < 8; ++q) s[q] = nw[q];
    }
    pin(s[1] + s[5] + dt);
    c1 = __builtin_readcyclecounter();
    w1 = __builtin_amdgcn_s_memrealtime();
    acc = s[1] + s[2] + s[3] + s[5];
  } else {
    // samples along a fixed state: primitives and fractions vary per
    // repetition so that nothing is hoisted out of the loop
    ImageParams ip;
    ip.fluid_gamma = 13.0 / 9;
    ip.r_high = 40.0;
    ip.Ne_unit = 4.3e5;
    ip.B_unit = 21.0;
    ip.L_unit = 9.2e14;
    ip.nu_obs = 230e9;
    ip.fallback_cos = 0.5;
    ip.sigma_cut = 100.0;
    Stencil st;
    st.mb = 0;
    st.base = 0;
    // two sets of 8 corner records per lane in LDS, read every repetition
    // (as the render kernel reads them from memory), so that nothing derived
    // from the corners can be hoisted out of the loop
    __shared__ float4 corners[2][16][64];
    for (int set = 0; set < 2; ++set)
#pragma unroll
      for (int c8 = 0; c8 < 8; ++c8) {
        const float w = 1.0f + 0.01f * c8 + 0.001f * lane + 0.003f * set;
        corners[set][2 * c8][lane] = make_float4(w, 0.02f * w, 0.01f, 0.2f);
        corners[set][2 * c8 + 1][lane] = make_float4(0.01f, 0.03f, 0.05f * w, 0.02f);
      }
    __syncthreads();
    Corners cr;
    double I = 0.0, trans = 1.0;
    pin(s[1]);
    c0 = __builtin_readcyclecounter();
    w0 = __builtin_amdgcn_s_memrealtime();
    for (int k = 0; k < steps; ++k) {
      const double f = 0.25 + 1e-5 * k;
#pragma unroll
      for (int c8 = 0; c8 < 8; ++c8) {
        cr.q[c8][0] = corners[k & 1][2 * c8][lane];
        cr.q[c8][1] = corners[k & 1][2 * c8 + 1][lane];
      }
      st.fr[0] = f;
      st.fr[1] = 0.5 - 0.5 * f;
      st.fr[2] = 0.75 * f;
      if (variant == 4 || variant == 6) {
        double P[8], j = 0.0, al = 0.0;
        interp_raw<double>(st, cr, P);
        if (variant == 6) {
          emission_from_prims<double>(ip, P, s, a, j, al);
          I += trans * (j * 1e-3);
          trans *= (1.0 - al * 1e-3);
        } else {
          I += P[0] + P[3] + P[7];
        }
      } else {
        float P[8], j = 0.0f, al = 0.0f;
        interp_raw<float>(st, cr, P);
        if (variant == 7) {
          emission_from_prims<float>(ip, P, s, a, j, al);
          I += trans * ((double)j * 1e-3);
          trans *= (1.0 - (double)al * 1e-3);
        } else {
          I += (double)(P[0] + P[3] + P[7]);
        }
      }
    }
    pin(I + trans);
    c1 = __builtin_readcyclecounter();
    w1 = __builtin_amdgcn_s_memrealtime();
    acc = I + trans;
  }
  out[lane] = acc;
  if (lane == 0) {
    t[0] = c1 - c0;
    t[1] = w1 - w0;
  }
}

}  // namespace mk

using namespace mk;

extern "C" int mk_step_cycles(int variant, int steps, double bhspin, double* h_out) {
  MK_REQUIRE(variant >= 0 && variant <= 7 && steps >= 1 && h_out && std::fabs(bhspin) < 1.0,
             "mk_step_cycles: bad arguments");
  double* d_out = nullptr;
  unsigned long long* d_t = nullptr;
  hipError_t e = hipMalloc(&d_out, 64 * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&d_t, 2 * sizeof(unsigned long long));
  unsigned long long t[2] = {0, 0};
  if (e == hipSuccess) {
    k_step_cycles<<<1, 64>>>(variant, steps, bhspin, d_out, d_t);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(t, d_t, sizeof t, hipMemcpyDeviceToHost);
  (void)hipFree(d_out);
  (void)hipFree(d_t);
  if (e != hipSuccess) {
    set_error("mk_step_cycles: %s", hipGetErrorString(e));
    return MK_ERR_HIP;
  }
  h_out[0] = (double)t[0] / steps;          // shader cycles per repetition
  h_out[1] = (double)t[1] * 10.0 / steps;   // ns per repetition (100 MHz counter)
  h_out[2] = t[1] ? (double)t[0] / ((double)t[1] * 10.0) : 0.0;  // cycles per ns (GHz)
  return MK_OK;
}
