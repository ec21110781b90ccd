// Cached per-axis tables computed on the host in FP64 with the reference's
// operation order, then uploaded once per (device, parameters).
#pragma once

#include "rsim_internal.hpp"

namespace rsim {

// (fftfreq(n)/spacing)^2 (big-wave/propagation.py:293-295)
struct AxisK2 {
  double* dev;
};
int axis_k2(int n, double spacing, const AxisK2** out);

// Field axis -> voxel axis map of Sample._apply_2d
// (big-wave/optical_element.py:538-546 and 555-563).
struct AxisMapDev {
  const int* i0;          // floor(idx) clipped to [0, len-2]
  const double* frac;     // idx - floor(idx) (unclipped floor)
  const float* fracf;     // frac rounded to FP32
  const uint8_t* inside;  // 0 <= idx < len - 1
};
struct AxisMap {
  AxisMapDev dev;
  bool all_integral;      // every inside idx within 1e-8 of an integer
  bool any_inside;
  // when all_integral: nearest voxel = a*i + c on the inside range [lo, hi)
  // (an integral map of a uniform grid is affine); affine == false if that
  // check fails
  bool affine;
  int a, c, lo, hi;
};
int axis_map(int n, double spacing, double scale, double pos, int grid_len, double pixel,
             const AxisMap** out);

int volume_db_dev(const rsim_volume* vol, const double** out);
int release_tables();

}  // namespace rsim
