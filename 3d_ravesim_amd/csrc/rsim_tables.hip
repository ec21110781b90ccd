#include <cmath>
#include <cstring>
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

std::mutex g_mu;
std::map<std::tuple<int, int, uint64_t>, AxisK2> g_k2;
std::map<std::tuple<int, int, uint64_t, uint64_t, uint64_t, int, uint64_t>, AxisMap> g_maps;
std::vector<void*> g_allocs;

}  // namespace

int axis_k2(int n, double spacing, const AxisK2** out) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto key = std::make_tuple(current_device(), n, bits(spacing));
  auto it = g_k2.find(key);
  if (it != g_k2.end()) {
    *out = &it->second;
    return RSIM_OK;
  }
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

int axis_map(int n, double spacing, double scale, double pos, int grid_len, double pixel,
             const AxisMap** out) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto key = std::make_tuple(current_device(), n, bits(spacing), bits(scale), bits(pos), grid_len,
                             bits(pixel));
  auto it = g_maps.find(key);
  if (it != g_maps.end()) {
    *out = &it->second;
    return RSIM_OK;
  }
  std::vector<int> i0(n), inear(n);
  std::vector<double> fr(n);
  std::vector<float> frf(n);
  std::vector<uint8_t> ins(n);
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
  AxisMap m;
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
  // one allocation and one copy: frac (8 n), i0 (4 n), fracf (4 n), inside (n)
  const size_t bytes = n * (4 + 8 + 4 + 1) + 64;
  std::vector<unsigned char> host(bytes);
  size_t o = 0;
  std::memcpy(host.data() + o, fr.data(), 8 * (size_t)n);
  const size_t o_i0 = (o += 8 * (size_t)n);
  std::memcpy(host.data() + o, i0.data(), 4 * (size_t)n);
  const size_t o_ff = (o += 4 * (size_t)n);
  std::memcpy(host.data() + o, frf.data(), 4 * (size_t)n);
  const size_t o_in = (o += 4 * (size_t)n);
  std::memcpy(host.data() + o, ins.data(), (size_t)n);
  unsigned char* d = nullptr;
  RSIM_HIP(hipMalloc((void**)&d, bytes));
  g_allocs.push_back(d);
  RSIM_HIP(hipMemcpy(d, host.data(), bytes, hipMemcpyHostToDevice));
  m.dev.frac = (const double*)d;
  m.dev.i0 = (const int*)(d + o_i0);
  m.dev.fracf = (const float*)(d + o_ff);
  m.dev.inside = (const uint8_t*)(d + o_in);
  auto r = g_maps.emplace(key, m);
  *out = &r.first->second;
  return RSIM_OK;
}

int volume_db_dev(const rsim_volume* vol, const double** out) {
  *out = vol->db_dev;
  return RSIM_OK;
}

int release_tables() {
  std::lock_guard<std::mutex> lk(g_mu);
  for (void* p : g_allocs) hipFree(p);
  g_allocs.clear();
  g_k2.clear();
  g_maps.clear();
  return RSIM_OK;
}

}  // namespace rsim
