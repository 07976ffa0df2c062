// mk_render.h -- device-side parameters of the fused render kernel.
#pragma once
#include <vector>

#include "mk_device.h"

namespace mk {

struct RenderDev {
  double a, ci, si, dist, lo, step, rEH, div, inv_div, tol;
  int res, N, tiles_x;
  long ntiles;
  ImageParams ip;
  double fp64_r, ring_r, ring_drdl;
  int spin_limit;     // polls before a bounded wait gives up (paired and triple forms)
  const int* order;   // tiles in the order they are handed out (nullptr: row-major)
  int* next_tile;     // classic form: tiles taken from this counter by each wave (nullptr: one tile per wave)
};

int fill_render_dev(const mk_render_params* p, RenderDev* rp);

// The 8x8 tiles of a res x res image (row-major index) ordered by the distance
// of the tile centre from the critical curve of a distant camera (the photon
// ring, whose rays take the most steps), nearest first.  Handing the longest
// tiles out first shortens a launch's tail; only the schedule changes.
std::vector<int> ring_first_tile_order(int res, double lo, double step, double a, double inc_deg);

// LDS budget for the staged block geometry + lookup grid (bytes)
constexpr size_t kMaxTableLds = 64 * 1024;
size_t table_bytes(const SnapView& sv);

}  // namespace mk
