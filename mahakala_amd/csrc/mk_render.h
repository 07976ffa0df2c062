// mk_render.h -- device-side parameters of the fused render kernel.
#pragma once
#include "mk_device.h"

namespace mk {

struct RenderDev {
  double a, ci, si, dist, lo, step, rEH, div, tol;
  int res, N, tiles_x;
  long ntiles;
  ImageParams ip;
  double fp64_r, ring_r, ring_drdl;
};

int fill_render_dev(const mk_render_params* p, RenderDev* rp);

}  // namespace mk
