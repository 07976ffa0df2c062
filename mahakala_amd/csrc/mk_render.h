// mk_render.h -- device-side parameters of the fused render kernel.
#pragma once
#include "mk_device.h"

namespace mk {

struct RenderDev {
  double a, ci, si, dist, lo, step, rEH, div, inv_div, tol;
  int res, N, tiles_x;
  long ntiles;
  ImageParams ip;
  double fp64_r, ring_r, ring_drdl;
  int spin_limit;  // polls before a bounded wait gives up (paired and triple forms)
};

int fill_render_dev(const mk_render_params* p, RenderDev* rp);

// LDS budget for the staged block geometry + lookup grid (bytes)
constexpr size_t kMaxTableLds = 64 * 1024;
size_t table_bytes(const SnapView& sv);

}  // namespace mk
