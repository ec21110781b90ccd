#include <atomic>
#include <cmath>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "rsim_tables.hpp"

// Host FP64 arithmetic follows the reference expression by expression.
#pragma clang fp contract(off)

namespace rsim {
namespace {

// |idx - nearest integer| (voxel units) below which an axis counts as integral
constexpr double kSnap = 1e-8;

uint64_t bits(double v) {
  uint64_t b;
  std::memcpy(&b, &v, 8);
  return b;
}

std::map<std::tuple<int, int, uint64_t>, AxisK2> g_k2;
std::map<std::tuple<int, int, uint64_t, uint64_t, uint64_t, int, uint64_t>, AxisMap> g_maps;
std::vector<void*> g_allocs;   // axis_k2 tables (a few per grid; freed on release)

// Registry of the evictable device tables of every cache: bytes, graph pins,
// last use, and how to drop the entry from its owning map.
struct Slot {
  int device;
  size_t bytes;
  int pins;
  unsigned long long tick;
  std::function<void()> drop;   // erase from the owning map (hipFree follows)
};
std::map<const void*, Slot> g_slots;
unsigned long long g_tick = 0;
std::atomic<unsigned long long> g_generation{1};

}  // namespace

std::recursive_mutex& table_mu() {
  static std::recursive_mutex mu;
  return mu;
}

unsigned long long cache_generation() { return g_generation.load(); }
void bump_cache_generation() { g_generation.fetch_add(1); }

namespace {
thread_local std::vector<const void*>* t_use_log = nullptr;
}

void table_use_log(std::vector<const void*>* log) { t_use_log = log; }

void table_touch(const void* dev) {
  auto it = g_slots.find(dev);
  if (it != g_slots.end()) it->second.tick = ++g_tick;
  if (t_use_log) t_use_log->push_back(dev);
}

int table_insert(void* dev, size_t bytes, std::function<void()> drop) {
  const int device = current_device();
  size_t held = 0;
  for (auto& kv : g_slots)
    if (kv.second.device == device) held += kv.second.bytes;
  // evict least recently used, unpinned entries of this device until it fits
  while (held + bytes > kTableCacheBytes) {
    auto victim = g_slots.end();
    for (auto it = g_slots.begin(); it != g_slots.end(); ++it)
      if (it->second.device == device && it->second.pins == 0 &&
          (victim == g_slots.end() || it->second.tick < victim->second.tick))
        victim = it;
    if (victim == g_slots.end()) break;   // everything left is pinned
    void* p = const_cast<void*>(victim->first);
    held -= victim->second.bytes;
    victim->second.drop();
    g_slots.erase(victim);
    RSIM_HIP(hipFree(p));
  }
  g_slots[dev] = Slot{device, bytes, 0, ++g_tick, std::move(drop)};
  if (t_use_log) t_use_log->push_back(dev);
  return RSIM_OK;
}

bool table_pin(const void* dev, int delta) {
  std::lock_guard<std::recursive_mutex> lk(table_mu());
  auto it = g_slots.find(dev);
  if (it == g_slots.end()) return false;
  it->second.pins += delta;
  if (it->second.pins < 0) it->second.pins = 0;
  return true;
}

void table_cache_stats(size_t* bytes, int* entries) {
  std::lock_guard<std::recursive_mutex> lk(table_mu());
  size_t b = 0;
  for (auto& kv : g_slots) b += kv.second.bytes;
  if (bytes) *bytes = b;
  if (entries) *entries = (int)g_slots.size();
}

void table_forget_all() {
  g_slots.clear();
}

int axis_k2(int n, double spacing, const AxisK2** out) {
  std::lock_guard<std::recursive_mutex> lk(table_mu());
  auto key = std::make_tuple(current_device(), n, bits(spacing));
  auto it = g_k2.find(key);
  if (it != g_k2.end()) {
    *out = &it->second;
    return RSIM_OK;
  }
  RSIM_CHECK(capture_miss("frequency tables"));
  std::vector<double> h(n);
  const double val = 1.0 / (double)n;   // numpy fftfreq: results * (1/(n*d))
  const int half = (n - 1) / 2 + 1;
  for (int i = 0; i < n; ++i) {
    const long long b = i < half ? i : (long long)i - n;
    const double f = ((double)b * val) / spacing;
    h[i] = f * f;
  }
  AxisK2 t;
  RSIM_HIP(hipMalloc((void**)&t.dev, sizeof(double) * n));
  RSIM_HIP(hipMemcpy(t.dev, h.data(), sizeof(double) * n, hipMemcpyHostToDevice));
  g_allocs.push_back(t.dev);
  auto r = g_k2.emplace(key, t);
  *out = &r.first->second;
  return RSIM_OK;
}

// device image: frac (8 n), i0 (4 n), fracf (4 n), inside (n), padding
size_t axis_map_bytes(int n) { return (size_t)n * (8 + 4 + 4 + 1) + 64; }

void axis_map_bind(AxisMap* m, const unsigned char* d, int n) {
  m->dev.frac = (const double*)d;
  m->dev.i0 = (const int*)(d + 8 * (size_t)n);
  m->dev.fracf = (const float*)(d + 12 * (size_t)n);
  m->dev.inside = (const uint8_t*)(d + 16 * (size_t)n);
}

void axis_map_build(int n, double spacing, double scale, double pos, int grid_len, double pixel,
                    AxisMap* mp, unsigned char* img) {
  double* fr = (double*)img;
  int* i0 = (int*)(img + 8 * (size_t)n);
  float* frf = (float*)(img + 12 * (size_t)n);
  uint8_t* ins = img + 16 * (size_t)n;
  std::vector<int> inear(n);
  bool integral = true, any = false;
  for (int i = 0; i < n; ++i) {
    // sim = (i - n/2.0)*spacing; idx = (sim*scale + pos + len*pixel*0.5)/pixel
    const double sim = ((double)i - n / 2.0) * spacing;
    const double idx = (sim * scale + pos + (double)grid_len * pixel * 0.5) / pixel;
    const double fl = std::floor(idx);
    const long long fraw = (long long)fl;
    fr[i] = idx - fl;
    frf[i] = (float)fr[i];
    ins[i] = (idx >= 0.0 && idx < (double)grid_len - 1.0) ? 1 : 0;
    long long c = fraw < 0 ? 0 : fraw;
    if (c > grid_len - 2) c = grid_len - 2;
    i0[i] = (int)c;
    // Nearest voxel.  Equal pitches leave rounding residues of ~1e-12 voxel
    // in idx (e.g. 1202 of 4096 positions at 0.5 um); within kSnap the
    // bilinear weight of the second corner is below 1e-8, far under FP32
    // resolution, so the axis is treated as integral.
    const double r = std::nearbyint(idx);
    inear[i] = 0;
    if (ins[i]) {
      inear[i] = (int)r;
      any = true;
      if (std::fabs(idx - r) > kSnap) integral = false;
    }
  }
  AxisMap& m = *mp;
  m.all_integral = integral;
  m.any_inside = any;
  m.affine = false;
  m.a = m.c = m.lo = m.hi = 0;
  if (integral) {
    int lo = -1, hi = -1;
    for (int i = 0; i < n; ++i)
      if (ins[i]) {
        if (lo < 0) lo = i;
        hi = i + 1;
      }
    bool ok = true;
    if (lo < 0) {
      lo = hi = 0;  // nothing inside: empty range
    } else {
      for (int i = lo; i < hi; ++i) ok = ok && ins[i];
      const int a = hi - lo > 1 ? inear[lo + 1] - inear[lo] : 1;
      const int c = inear[lo] - a * lo;
      for (int i = lo; i < hi && ok; ++i) ok = inear[i] == a * i + c && inear[i] < grid_len;
      m.a = a;
      m.c = c;
    }
    m.affine = ok;
    m.lo = lo;
    m.hi = hi;
  }
}

int axis_map(int n, double spacing, double scale, double pos, int grid_len, double pixel,
             const AxisMap** out) {
  std::lock_guard<std::recursive_mutex> lk(table_mu());
  auto key = std::make_tuple(current_device(), n, bits(spacing), bits(scale), bits(pos), grid_len,
                             bits(pixel));
  auto it = g_maps.find(key);
  if (it != g_maps.end()) {
    table_touch(it->second.dev.frac);
    *out = &it->second;
    return RSIM_OK;
  }
  RSIM_CHECK(capture_miss("voxel axis maps"));
  const size_t bytes = axis_map_bytes(n);
  std::vector<unsigned char> host(bytes, 0);
  AxisMap m;
  axis_map_build(n, spacing, scale, pos, grid_len, pixel, &m, host.data());
  unsigned char* d = nullptr;
  RSIM_HIP(hipMalloc((void**)&d, bytes));
  RSIM_HIP(hipMemcpy(d, host.data(), bytes, hipMemcpyHostToDevice));
  axis_map_bind(&m, d, n);
  RSIM_CHECK(table_insert(d, bytes, [key] { g_maps.erase(key); }));
  auto r = g_maps.emplace(key, m);
  *out = &r.first->second;
  return RSIM_OK;
}

int volume_db_dev(const rsim_volume* vol, const double** out) {
  *out = vol->db_dev;
  return RSIM_OK;
}

int release_tables() {
  std::lock_guard<std::recursive_mutex> lk(table_mu());
  for (void* p : g_allocs) hipFree(p);
  g_allocs.clear();
  g_k2.clear();
  for (auto& kv : g_maps) hipFree(const_cast<double*>(kv.second.dev.frac));
  g_maps.clear();
  return RSIM_OK;
}

}  // namespace rsim
