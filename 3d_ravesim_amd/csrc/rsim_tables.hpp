// Cached per-axis tables computed on the host in FP64 with the reference's
// operation order, then uploaded once per (device, parameters).
#pragma once

#include <functional>
#include <mutex>
#include <vector>

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
// Host half of axis_map: the metadata of the map into *m (device pointers
// unset) and its device image, axis_map_bytes(n) bytes, into img.
size_t axis_map_bytes(int n);
void axis_map_build(int n, double spacing, double scale, double pos, int grid_len, double pixel,
                    AxisMap* m, unsigned char* img);
// point m's device pointers at an image written by axis_map_build
void axis_map_bind(AxisMap* m, const unsigned char* dev, int n);

// ---- bounds of the device table caches ----
// Every cache of device tables (axis maps, chirp tables) holds at most
// kTableCacheBytes per device; inserting beyond that frees the least recently
// used entries that no captured graph pins (hipFree, which waits for the
// device).  Cone-beam runs of the fused engine keep their per-slice tables
// in a buffer of their own, freed at the end of the run.
constexpr size_t kTableCacheBytes = 256ull << 20;
// Pin / unpin the cached table whose device image starts at dev (captured
// graphs hold pins for their life); false when no cache holds it.
bool table_pin(const void* dev, int delta);
// Bytes and entries held by the device table caches (all devices).
void table_cache_stats(size_t* bytes, int* entries);
// Incremented by rsim_release_caches: objects that captured cached pointers
// compare it with the value they saw.
unsigned long long cache_generation();
void bump_cache_generation();
// Shared by the caches (one lock: an insertion may evict from another cache).
std::recursive_mutex& table_mu();
// Under table_mu: register a new cached device allocation (evicting others
// first if the cache is full; drop erases the entry from its owning map),
// mark one as used, forget every registration (on release).
int table_insert(void* dev, size_t bytes, std::function<void()> drop);
void table_touch(const void* dev);
void table_forget_all();
// While set (per thread), every cached table looked up or created is
// appended to *log (graph captures pin what they read).
void table_use_log(std::vector<const void*>* log);

int volume_db_dev(const rsim_volume* vol, const double** out);
int release_tables();

}  // namespace rsim
