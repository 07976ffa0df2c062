// mk_internal.h -- error plumbing and launch helpers shared by the sources.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>

#include "../../include/mahakala_hip.h"

namespace mk {

void set_error(const char* fmt, ...);

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

inline unsigned grid_for(long n, int block) { return (unsigned)((n + block - 1) / block); }

// mk_shadow_bisect on `st` with the caller's device staging d_stage (3 n
// doubles) and host radii out; returns after the radii are in h_radii.  The
// same radii, bit for bit.
hipError_t shadow_bisect_on_stream(double a, double inc, const double* h_angles, long n, int max_steps,
                                   double error_allowed, int max_it, double dist, double* d_stage, double* h_radii,
                                   hipStream_t st);

}  // namespace mk

#define MK_HIP(call)                                                                    \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess) {                                                             \
      mk::set_error("%s:%d %s: %s", __FILE__, __LINE__, #call, hipGetErrorString(e_)); \
      return MK_ERR_HIP;                                                                \
    }                                                                                   \
  } while (0)

#define MK_LAUNCH_CHECK()                                                                   \
  do {                                                                                      \
    hipError_t e_ = hipGetLastError();                                                      \
    if (e_ != hipSuccess) {                                                                 \
      mk::set_error("%s:%d kernel launch: %s", __FILE__, __LINE__, hipGetErrorString(e_)); \
      return MK_ERR_HIP;                                                                    \
    }                                                                                       \
  } while (0)

#define MK_REQUIRE(cond, ...)      \
  do {                             \
    if (!(cond)) {                 \
      mk::set_error(__VA_ARGS__);  \
      return MK_ERR_ARG;           \
    }                              \
  } while (0)
