// mk_snapshot.h -- host-side handle of a device-resident snapshot.
#pragma once
#include <hip/hip_runtime.h>

#include "mk_device.h"

struct mk_snapshot {
  long nmb = 0;
  int n = 0;                  // cells per block side (halo excluded)
  float4* d_cells = nullptr;  // nmb * (n+2)^3 * 2 : 8-prim records
  double* d_meta = nullptr;   // nmb * 16 : lo[3], hi[3], v0[3], dx[3]
  int32_t* d_grid = nullptr;  // G[2] * G[1] * G[0] leaf ids
  int G[3] = {0, 0, 0};
  double g0[3] = {0, 0, 0}, ginv[3] = {0, 0, 0}, dom_lo[3] = {0, 0, 0}, dom_hi[3] = {0, 0, 0};
  size_t bytes = 0;
  mk::SnapView view() const;
};
