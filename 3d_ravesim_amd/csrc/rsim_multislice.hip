// Multislice engines for Sample._apply_2d (big-wave/optical_element.py:502-596).
//
// PLAIN: per slice a transmission kernel, FFT2, chirp+cutoff, IFFT2 -- the
// reference's unfused sequence (fast-wave simulation.cpp:681-714), in the
// field precision.
//
// FUSED (complex64): two HBM passes per slice-step.
//   row pass    V = FFT_x( T_s * IFFT_x(W) )   (first slice: FFT_x(T_s * u))
//   column pass W = IFFT_y( M * FFT_y(V) ),  M = H(kx,ky) * cutoff / (nx*ny)
// and a final row IFFT.  The separable chirp exp(i(theta_g+theta_x+theta_y))
// comes from FP64 axis tables (phases reduced mod 2pi in FP64) and is applied
// to the FP32 field in FP64 so its rounding does not repeat coherently from
// slice to slice; voxel transmission factors are FP64 values split into FP32
// hi+lo pairs for the same reason.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <thread>
#include <mutex>
#include <tuple>
#include <vector>

#include "rsim_fft_cfg.hpp"
#include "rsim_internal.hpp"
#include "rsim_tables.hpp"
#include "rsim_fused.hpp"

// Captured graph of a batched run.
struct rsim_batch {
  int device = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int groups = 0;
  int kernels = 0;
  unsigned long long generation = 0;   // cache generation the graph's pointers belong to
  std::vector<const void*> pinned;     // cached tables the graph reads
};

namespace rsim {

int transmission(void* u, int nx, int ny, int dtype, double dx, double dy, double wl,
                 const rsim_volume* vol, int slice, double px, double py, double pz, double xpos,
                 double ypos, double coord_scale, hipStream_t s);

namespace {

template <typename TW>
bool fused_row_b(int N, int v, int mode, int tm, const RowBatch& b, int count, int ny,
                 const StageTw<TW>& tw, hipStream_t s, const AdjArgs& aj = AdjArgs{}) {
  switch (N) {
    case 256: return row_go<256, TW>(v, mode, tm, b, count, ny, tw, s, aj);
    case 512: return row_go<512, TW>(v, mode, tm, b, count, ny, tw, s, aj);
    case 1024: return row_go<1024, TW>(v, mode, tm, b, count, ny, tw, s, aj);
    case 2048: return row_go<2048, TW>(v, mode, tm, b, count, ny, tw, s, aj);
    case 3072: return row_go<3072, TW>(v, mode, tm, b, count, ny, tw, s, aj);
    case 4096: return row_go<4096, TW>(v, mode, tm, b, count, ny, tw, s, aj);
    case 6144: return row_go<6144, TW>(v, mode, tm, b, count, ny, tw, s, aj);
    case 8192: return row_go<8192, TW>(v, mode, tm, b, count, ny, tw, s, aj);
    case 16384: return row_go<16384, TW>(v, mode, tm, b, count, ny, tw, s, aj);
    default: return false;
  }
}

template <typename TW>
bool fused_col_b(int N, int v, bool spec_out, const ColBatch& b, int count, int nx,
                 const StageTw<TW>& tw, float ss, hipStream_t s) {
  switch (N) {
    case 256: return col_go<256, TW>(v, spec_out, b, count, nx, tw, ss, s);
    case 512: return col_go<512, TW>(v, spec_out, b, count, nx, tw, ss, s);
    case 1024: return col_go<1024, TW>(v, spec_out, b, count, nx, tw, ss, s);
    case 2048: return col_go<2048, TW>(v, spec_out, b, count, nx, tw, ss, s);
    case 3072: return col_go<3072, TW>(v, spec_out, b, count, nx, tw, ss, s);
    case 4096: return col_go<4096, TW>(v, spec_out, b, count, nx, tw, ss, s);
    case 6144: return col_go<6144, TW>(v, spec_out, b, count, nx, tw, ss, s);
    case 8192: return col_go<8192, TW>(v, spec_out, b, count, nx, tw, ss, s);
    default: return false;
  }
}

// Columns per workgroup of column variant v of length N (0: no such variant).
int fused_col_width(int N, int v) {
  switch (N) {
    case 256: return col_width<256>(v);
    case 512: return col_width<512>(v);
    case 1024: return col_width<1024>(v);
    case 2048: return col_width<2048>(v);
    case 3072: return col_width<3072>(v);
    case 4096: return col_width<4096>(v);
    case 6144: return col_width<6144>(v);
    case 8192: return col_width<8192>(v);
    default: return 0;
  }
}

// single-field launches; in_tile / out_tile: tiled intermediate layout
template <typename TW>
bool fused_row(int N, int v, int mode, int tm, const void* in, void* out, int ny,
               const StageTw<TW>& tw, const TransArgsF& ta, hipStream_t s,
               const AdjArgs& aj = AdjArgs{}, int in_tile = 0) {
  RowBatch b{};
  b.inst[0].in = reinterpret_cast<const cpx<float>*>(in);
  b.inst[0].out = reinterpret_cast<cpx<float>*>(out);
  b.inst[0].ta = ta;
  b.in_tile = in_tile;
  return fused_row_b<TW>(N, v, mode, tm, b, 1, ny, tw, s, aj);
}

template <typename TW>
bool fused_col(int N, int v, bool spec_out, const void* in, void* out, void* spec, int nx,
               const StageTw<TW>& tw, const double2* Bx, const double2* Ay, const int* Rmax,
               float ss, hipStream_t s, int out_tile = 0) {
  ColBatch b{};
  b.inst[0] = ColInst{reinterpret_cast<const cpx<float>*>(in), reinterpret_cast<cpx<float>*>(out),
                      reinterpret_cast<cpx<float>*>(spec), Bx, Ay, Rmax};
  b.out_tile = out_tile;
  return fused_col_b<TW>(N, v, spec_out, b, 1, nx, tw, ss, s);
}

// FP64 chirp tables of the fused column pass.
struct ChirpTables {
  double2* Bx = nullptr;   // g * exp(i theta_x) / (nx*ny)
  double2* Ay = nullptr;   // exp(i theta_y)
  int* Rmax = nullptr;     // largest |ky index| kept for each column, -1: none
};

std::map<std::tuple<int, int, int, uint64_t, uint64_t, uint64_t, uint64_t, uint64_t, uint64_t, uint64_t>,
         ChirpTables>
    g_chirp;

uint64_t bits(double v) {
  uint64_t b;
  std::memcpy(&b, &v, 8);
  return b;
}

#pragma clang fp contract(off)
double reduce_phase(double th) {
  const double two_pi = 6.283185307179586476925286766559;
  return th - two_pi * std::nearbyint(th / two_pi);
}

std::vector<double> host_k2(int n, double spacing) {
  std::vector<double> h(n);
  const double val = 1.0 / (double)n;
  const int half = (n - 1) / 2 + 1;
  for (int i = 0; i < n; ++i) {
    const long long b = i < half ? i : (long long)i - n;
    const double f = ((double)b * val) / spacing;
    h[i] = f * f;
  }
  return h;
}

// Device image of one chirp table set: Bx (16 nx), Ay (16 ny), Rmax (4 nx).
size_t chirp_bytes(int nx, int ny) { return 16 * (size_t)nx + 16 * (size_t)ny + 4 * (size_t)nx; }

void chirp_bind(ChirpTables* t, unsigned char* dev, int nx, int ny) {
  t->Bx = reinterpret_cast<double2*>(dev);
  t->Ay = reinterpret_cast<double2*>(dev + 16 * (size_t)nx);
  t->Rmax = reinterpret_cast<int*>(dev + 16 * (size_t)(nx + ny));
}

// Global phase exp(2 pi i dz / wl) of propagation over dz
// (big-wave/propagation.py:428) as a unit complex number: libm's cos and sin
// reduce the argument (1e10 rad and more at metre distances) exactly, where a
// reduction by a double-precision 2 pi would be off by up to an ulp of it.
double2 global_phase(double wl, double dz) {
  const double g = ((2.0 * M_PI) / wl) * dz;
  return make_double2(std::cos(g), std::sin(g));
}

// Host FP64 build of the chirp tables of propagation over dz into img, with
// global phase gp (global_phase(wl, dz), or the product of several steps'
// phases when propagations are folded into one).
void chirp_build(int nx, int ny, double dx, double dy, double wl, double dz, double freq,
                 double2 gp, unsigned char* img) {
  const std::vector<double> kx2 = host_k2(nx, dx), ky2 = host_k2(ny, dy);
  const double ctheta = ((-M_PI) * wl) * dz;
  const double inv_n = 1.0 / ((double)nx * (double)ny);
  double* bx = reinterpret_cast<double*>(img);
  double* ay = reinterpret_cast<double*>(img + 16 * (size_t)nx);
  int* rmax = reinterpret_cast<int*>(img + 16 * (size_t)(nx + ny));
  for (int x = 0; x < nx; ++x) {
    const double th = reduce_phase(ctheta * kx2[x]);
    const double c = std::cos(th), sn = std::sin(th);
    bx[2 * x] = (gp.x * c - gp.y * sn) * inv_n;
    bx[2 * x + 1] = (gp.x * sn + gp.y * c) * inv_n;
  }
  for (int y = 0; y < ny; ++y) {
    const double th = reduce_phase(ctheta * ky2[y]);
    ay[2 * y] = std::cos(th);
    ay[2 * y + 1] = std::sin(th);
  }
  const double fsq = 2.0 * (freq * freq);
  // ky2 by |index|: index m and -m have equal squares; -ny/2 exists for even
  // ny.  ky2[m] is non-decreasing in m for m <= ny/2, and so is the rounded
  // sum with kx2[x]: the first m that leaves the disc is a binary search.
  const int mmax = ny / 2;
  for (int x = 0; x < nx; ++x) {
    int lo = 0, hi = mmax + 1;   // first m in [0, mmax + 1) with ky2[m] + kx2[x] > fsq
    while (lo < hi) {
      const int mid = (lo + hi) / 2;
      if (ky2[mid] + kx2[x] > fsq) hi = mid;
      else lo = mid + 1;
    }
    rmax[x] = lo - 1;
  }
}

int chirp_tables(int nx, int ny, double dx, double dy, double wl, double dz, double freq,
                 const ChirpTables** out, const double2* gphase = nullptr) {
  std::lock_guard<std::recursive_mutex> lk(table_mu());
  const double2 gp = gphase ? *gphase : global_phase(wl, dz);
  auto key = std::make_tuple(current_device(), nx, ny, bits(dx), bits(dy), bits(wl), bits(dz),
                             bits(freq), bits(gp.x), bits(gp.y));
  auto it = g_chirp.find(key);
  if (it != g_chirp.end()) {
    table_touch(it->second.Bx);
    *out = &it->second;
    return RSIM_OK;
  }
  RSIM_CHECK(capture_miss("chirp tables"));
  const size_t nb = chirp_bytes(nx, ny);
  std::vector<unsigned char> host(nb);
  chirp_build(nx, ny, dx, dy, wl, dz, freq, gp, host.data());
  ChirpTables t;
  unsigned char* dev = nullptr;
  RSIM_HIP(hipMalloc((void**)&dev, nb));
  RSIM_HIP(hipMemcpy(dev, host.data(), nb, hipMemcpyHostToDevice));
  chirp_bind(&t, dev, nx, ny);
  RSIM_CHECK(table_insert(dev, nb, [key] { g_chirp.erase(key); }));
  auto r = g_chirp.emplace(key, t);
  *out = &r.first->second;
  return RSIM_OK;
}

// Per-call material tables of the fused row pass (tiny; synchronous upload
// into a per-device buffer reused across calls).
struct MatBuf {
  void* dev = nullptr;
};
// one table per (device, volume, K): runs in flight on other streams, and the
// instances of a batch, each keep their own
std::map<std::tuple<int, unsigned long long, uint64_t>, MatBuf> g_mat;

int material_tables(const rsim_volume* vol, double K, const float4** lut, const float2** dbf) {
  std::lock_guard<std::recursive_mutex> lk(table_mu());
  const auto key = std::make_tuple(current_device(), vol->serial, bits(K));
  auto it = g_mat.find(key);
  if (it != g_mat.end()) {
    *lut = (const float4*)it->second.dev;
    *dbf = (const float2*)((char*)it->second.dev + 256 * 16);
    return RSIM_OK;
  }
  RSIM_CHECK(capture_miss("material tables"));
  MatBuf mb;
  RSIM_HIP(hipMalloc(&mb.dev, 256 * (16 + 8)));
  g_mat[key] = mb;
  std::vector<float4> l(256);
  std::vector<float2> d(256);
  for (int m = 0; m < 256; ++m) {
    double dr = 0, di = 0;
    if (m < vol->n_db) {
      dr = vol->db_host[2 * m];
      di = vol->db_host[2 * m + 1];
    }
    // exp(-2j*pi*t/wl*conj(db)) = exp(-K*beta) * exp(-i*K*delta)
    const double er = -(K * di), ei = -(K * dr);
    const double e = std::exp(er);
    const double tr = e * std::cos(ei), ti = e * std::sin(ei);
    const float hr = (float)tr, hi = (float)ti;
    l[m] = make_float4(hr, hi, (float)(tr - (double)hr), (float)(ti - (double)hi));
    d[m] = make_float2((float)dr, (float)di);
  }
  RSIM_HIP(hipMemcpy(mb.dev, l.data(), 256 * 16, hipMemcpyHostToDevice));
  RSIM_HIP(hipMemcpy((char*)mb.dev + 256 * 16, d.data(), 256 * 8, hipMemcpyHostToDevice));
  *lut = (const float4*)mb.dev;
  *dbf = (const float2*)((char*)mb.dev + 256 * 16);
  return RSIM_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// Launch-shape autotuner: the first fused call for a (device, nx, ny,
// transmission mode) times every row-pass and column-pass variant on a
// scratch field (hipEvents, 1 warm-up + 3 timed launches each) and keeps the
// fastest.  RAVESIM_FUSED_VARIANT="row,col" overrides the choice.
// ---------------------------------------------------------------------------
namespace {
// Tuned launch shapes of one (device, nx, ny, transmission, arithmetic).
// (rv, cv): the fastest slice with both passes in place on one buffer.
// tile > 0: slices run with the column pass storing tiles of that width,
// which the next row pass reads (ping-pong over two buffers), with variants
// (trv, tcv): the fastest such pair, used unless an in-place pair is more
// than 1 % faster.
struct FusedChoice {
  int rv = 0, cv = 0;
  int trv = 0, tcv = 0, tile = 0;
};
std::map<std::tuple<int, int, int, int>, FusedChoice> g_tune;
std::mutex g_tune_mu;

// Timing field of the autotuner: hashed values in [-1, 1) (a field of zeros
// draws less power than real data and favours VALU-heavy shapes).
__global__ void tune_fill_kernel(float* p, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    uint32_t h = (uint32_t)i * 2654435761u;
    h ^= h >> 15;
    h *= 2246822519u;
    h ^= h >> 13;
    p[i] = (float)(h >> 8) * (1.0f / 8388608.0f) - 1.0f;
  }
}

int tune_fill(void* p, size_t bytes, hipStream_t s) {
  const long long n = (long long)(bytes / 4);
  const int blocks = (int)std::min<long long>((n + 255) / 256, 16384);
  hipLaunchKernelGGL(tune_fill_kernel, dim3(blocks), dim3(256), 0, s, (float*)p, n);
  RSIM_LAUNCHED();
  return RSIM_OK;
}

// Average time of one launch over reps launches after one untimed launch;
// -1 when the launch shape does not exist or fails.
int time_launches(hipStream_t s, const std::function<bool()>& launch, float* ms, int reps) {
  hipEvent_t a, b;
  RSIM_HIP(hipEventCreate(&a));
  RSIM_HIP(hipEventCreate(&b));
  bool ok = launch();
  hipError_t e = hipGetLastError();
  if (!ok || e != hipSuccess) {
    hipEventDestroy(a);
    hipEventDestroy(b);
    *ms = -1.f;
    return RSIM_OK;
  }
  RSIM_HIP(hipEventRecord(a, s));
  for (int i = 0; i < reps; ++i) launch();
  RSIM_HIP(hipEventRecord(b, s));
  RSIM_HIP(hipEventSynchronize(b));
  RSIM_HIP(hipEventElapsedTime(ms, a, b));
  *ms /= (float)reps;
  e = hipGetLastError();
  if (e != hipSuccess) *ms = -1.f;
  hipEventDestroy(a);
  hipEventDestroy(b);
  return RSIM_OK;
}

// Tile width usable between column variant cv and the next row pass, or 0.
int tile_width(int nx, int ny, int cv) {
  const int w = fused_col_width(ny, cv);
  return (w >= 4 && nx % w == 0) ? w : 0;
}

void pick(const FusedChoice& fc, int* rv, int* cv, int* tile) {
  if (tile && fc.tile) {
    *rv = fc.trv;
    *cv = fc.tcv;
    *tile = fc.tile;
    return;
  }
  *rv = fc.rv;
  *cv = fc.cv;
  if (tile) *tile = 0;
}
}  // namespace

// Launch shapes of a fused slice.  tile == nullptr: the caller runs both
// passes in place on one buffer; otherwise *tile receives the tile width of
// the intermediate layout when a tiled slice is faster (0: in place).
// RAVESIM_FUSED_VARIANT="row,col" or "row,col,tiled" (tiled 1: column tiles
// where the column width allows): a fixed schedule instead of the tuned one.
static bool env_schedule(int nx, int ny, int* rv, int* cv, int* tile) {
  const char* env = std::getenv("RAVESIM_FUSED_VARIANT");
  if (!env) return false;
  int r = 0, c = 0, t = 0;
  const int got = std::sscanf(env, "%d,%d,%d", &r, &c, &t);
  if (got < 2 || r < 0 || r >= kRowVariants || c < 0 || c >= kColVariants) return false;
  *rv = r;
  *cv = c;
  if (tile) *tile = (got == 3 && t) ? tile_width(nx, ny, c) : 0;
  return true;
}

// Three-pass column tables (ny = kBigCol): the 128-point stage tables and W_ny.
template <typename TW>
struct BigColTables {
  StageTw<TW> t128{};
  const BigTwElem<TW>* wN = nullptr;
};

template <typename TW>
int fused_schedule(int nx, int ny, int tm, const StageTw<TW>& tx, const StageTw<TW>& ty,
                   const TransArgsF& ta, const ChirpTables* ct, hipStream_t s, int* rv, int* cv,
                   int* tile = nullptr, const BigColTables<TW>* big = nullptr) {
  if (env_schedule(nx, ny, rv, cv, tile)) return RSIM_OK;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  const int math = std::is_same<TW, twc>::value ? 4 : sizeof(TW) == 8 ? 2 : 0;
  auto key = std::make_tuple(current_device(), nx, ny, tm + 4 * math);
  auto it = g_tune.find(key);
  if (it != g_tune.end()) {
    pick(it->second, rv, cv, tile);
    return RSIM_OK;
  }
  RSIM_CHECK(capture_miss("launch-shape tuning"));
  const size_t tbytes = (size_t)nx * ny * 8;
  struct Bufs {
    void* a = nullptr;
    void* b = nullptr;
    ~Bufs() {
      if (a) (void)hipFree(a);
      if (b) (void)hipFree(b);
    }
  } buf;
  RSIM_HIP(hipMalloc(&buf.a, tbytes));
  RSIM_HIP(hipMalloc(&buf.b, tbytes));
  void* tmp = buf.a;
  void* tmp2 = buf.b;
  const bool verbose = std::getenv("RAVESIM_TUNE_VERBOSE") != nullptr;
  const char* mname = math == 4 ? "comp32" : math ? "fp64" : "fp32";
  auto row = [&](int v, const void* in, void* out, int in_tile) {
    return fused_row(nx, v, ROW_MID, tm, in, out, ny, tx, ta, s, AdjArgs{}, in_tile);
  };
  auto col = [&](int v, const void* in, void* out, int out_tile) {
    return fused_col(ny, v, false, in, out, out, nx, ty, ct->Bx, ct->Ay, ct->Rmax, 1.f, s,
                     out_tile);
  };
  // each shape on its own, in place, two interleaved rounds (least of the two) ...
  float tr[kRowVariants], tc[kColVariants];
  for (float& x : tr) x = 1e30f;
  for (float& x : tc) x = 1e30f;
  for (int round = 0; round < 2; ++round) {
    for (int v = 0; v < kRowVariants; ++v) {
      float ms = -1.f;
      RSIM_CHECK(tune_fill(tmp, tbytes, s));
      RSIM_CHECK(time_launches(s, [&] { return row(v, tmp, tmp, 0); }, &ms, 5));
      if (ms > 0.f && ms < tr[v]) tr[v] = ms;
    }
    for (int v = 0; v < kColVariants; ++v) {
      float ms = -1.f;
      RSIM_CHECK(tune_fill(tmp, tbytes, s));
      RSIM_CHECK(time_launches(s, [&] { return col(v, tmp, tmp, 0); }, &ms, 5));
      if (ms > 0.f && ms < tc[v]) tc[v] = ms;
    }
  }
  auto top2 = [](const float* t, int n, int* out) {
    out[0] = out[1] = -1;
    for (int v = 0; v < n; ++v) {
      if (t[v] >= 1e30f) continue;
      if (out[0] < 0 || t[v] < t[out[0]]) { out[1] = out[0]; out[0] = v; }
      else if (out[1] < 0 || t[v] < t[out[1]]) out[1] = v;
    }
  };
  int rb[2], cb[2];
  top2(tr, kRowVariants, rb);
  top2(tc, kColVariants, cb);
  if (verbose) {
    for (int v = 0; v < kRowVariants; ++v)
      std::fprintf(stderr, "[rsim tune] %dx%d math=%s row variant %d: %.1f us\n", nx, ny, mname, v,
                   tr[v] < 1e30f ? tr[v] * 1e3f : -1.f);
    for (int v = 0; v < kColVariants; ++v)
      std::fprintf(stderr, "[rsim tune] %dx%d math=%s col variant %d: %.1f us\n", nx, ny, mname, v,
                   tc[v] < 1e30f ? tc[v] * 1e3f : -1.f);
  }
  // ... then the leading pairs as the slice loop runs them: row pass and
  // column pass in place on one buffer, and (where the column width allows)
  // the tiled pair, row pass from tiles in one buffer into the other and
  // column pass back as tiles.  The pair, not the sum of its parts, is what
  // a slice costs.
  FusedChoice fc;
  fc.rv = rb[0] < 0 ? 0 : rb[0];
  fc.cv = cb[0] < 0 ? 0 : cb[0];
  if (big) {
    // three-pass columns: each tile width as the slice runs it, all three
    // passes in place
    float bt = 1e30f;
    for (int v = 0; v < kBigColVariants; ++v) {
      float ms = -1.f;
      RSIM_CHECK(tune_fill(tmp, tbytes, s));
      RSIM_CHECK(time_launches(s, [&] {
        return big_col_go<TW>(v, tmp, nx, ny, big->t128, big->wN, ct->Bx, ct->Ay, ct->Rmax, s);
      }, &ms, 3));
      if (verbose)
        std::fprintf(stderr, "[rsim tune] %dx%d math=%s three-pass column variant %d: %.1f us\n",
                     nx, ny, mname, v, ms * 1e3f);
      if (ms > 0.f && ms < bt) {
        bt = ms;
        fc.cv = v;
      }
    }
  }
  float best = 1e30f, best_tiled = 1e30f;
  if (rb[0] >= 0 && cb[0] >= 0) {
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j) {
        if (rb[i] < 0 || cb[j] < 0) continue;
        const int w = tile_width(nx, ny, cb[j]);
        for (int tiled = 0; tiled < 2; ++tiled) {
          if (tiled && !w) continue;
          float ms = -1.f;
          RSIM_CHECK(tune_fill(tmp, tbytes, s));
          if (tiled) RSIM_CHECK(tune_fill(tmp2, tbytes, s));
          if (tiled)
            RSIM_CHECK(time_launches(
                s, [&] { return row(rb[i], tmp2, tmp, w) && col(cb[j], tmp, tmp2, w); }, &ms, 8));
          else
            RSIM_CHECK(time_launches(
                s, [&] { return row(rb[i], tmp, tmp, 0) && col(cb[j], tmp, tmp, 0); }, &ms, 8));
          if (verbose)
            std::fprintf(stderr, "[rsim tune] %dx%d math=%s pair (%d, %d)%s: %.1f us\n", nx, ny,
                         mname, rb[i], cb[j], tiled ? " tiled" : "", ms * 1e3f);
          if (ms <= 0.f) continue;
          if (!tiled && ms < best) {
            best = ms;
            fc.rv = rb[i];
            fc.cv = cb[j];
          }
          if (tiled && ms < best_tiled) {
            best_tiled = ms;
            fc.trv = rb[i];
            fc.tcv = cb[j];
            fc.tile = w;
          }
        }
      }
  }
  // The tiled chain moves fewer HBM bytes (no partial-line column stores);
  // it is kept unless the in-place chain is more than 1 % faster.
  if (fc.tile && !(best_tiled <= 1.01f * best)) fc.tile = 0;
  if (verbose)
    std::fprintf(stderr, "[rsim tune] %dx%d math=%s chosen (%d, %d)%s\n", nx, ny, mname,
                 fc.tile ? fc.trv : fc.rv, fc.tile ? fc.tcv : fc.cv, fc.tile ? " tiled" : "");
  RSIM_HIP(hipStreamSynchronize(s));
  g_tune[key] = fc;
  pick(fc, rv, cv, tile);
  return RSIM_OK;
}

int fused_schedule_lookup(int nx, int ny, int* rv, int* cv, int* tile) {
  if (env_schedule(nx, ny, rv, cv, tile)) return RSIM_OK;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  *rv = *cv = -1;
  if (tile) *tile = -1;
  // the material-id lookup in FP32 butterflies (the default voxel slice loop),
  // else the first tuned entry of the shape
  auto it = g_tune.find(std::make_tuple(current_device(), nx, ny, (int)TM_LUT));
  if (it != g_tune.end()) {
    pick(it->second, rv, cv, tile);
    return RSIM_OK;
  }
  for (auto& kv : g_tune) {
    if (std::get<0>(kv.first) == current_device() && std::get<1>(kv.first) == nx &&
        std::get<2>(kv.first) == ny) {
      pick(kv.second, rv, cv, tile);
      break;
    }
  }
  return RSIM_OK;
}

int release_tune() {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune.clear();
  return RSIM_OK;
}

// Material tables of a volume being destroyed.
int release_volume_tables(unsigned long long serial) {
  std::lock_guard<std::recursive_mutex> lk(table_mu());
  for (auto it = g_mat.begin(); it != g_mat.end();) {
    if (std::get<1>(it->first) == serial) {
      RSIM_HIP(hipFree(it->second.dev));
      it = g_mat.erase(it);
    } else {
      ++it;
    }
  }
  return RSIM_OK;
}

static void release_run_tables();
int release_multislice_tables() {
  std::lock_guard<std::recursive_mutex> lk(table_mu());
  for (auto& kv : g_chirp) hipFree(kv.second.Bx);   // one allocation per table set
  g_chirp.clear();
  release_run_tables();
  for (auto& kv : g_mat) hipFree(kv.second.dev);
  g_mat.clear();
  return RSIM_OK;
}

// Row lengths with a fused row pass; column lengths with a single-pass
// fused column kernel, or (kBigCol) the three-pass schedule.
constexpr int kBigCol = 16384;
// Default butterfly arithmetic.  FP32 stage twiddles have |W| != 1 by up to
// ~6e-8, and a spectral component meets the same twiddles in every slice, so
// FP32's error grows linearly with depth; compensated FP32 grows far slower.
// Detector error of FP32 per slice (relative L2 versus the complex128 engine,
// 0.5 m leg, area_v1 4 x 4; seeded sphere phantoms at 128 and 256 slices,
// MI355X), per transform length:
static double fp32_error_per_slice(int n) {
  switch (n) {
    case 256: return 1.94e-8;
    case 512: return 4.4e-8;
    case 1024: return 2.2e-8;
    case 2048: return 1.72e-8;
    case 3072: return 2.07e-8;
    case 4096: return 1.56e-8;
    case 6144: return 1.84e-8;
    case 8192: return 2.5e-8;
    case 16384: return 4.0e-8;   // grows faster than linearly beyond 128 slices
    default: return 4.4e-8;
  }
}
// FP32 is the default while its predicted detector error stays within half
// the 1e-5 bar; deeper volumes default to compensated FP32.
static int fp32_depth_limit(int nx, int ny) {
  const double c = 0.5 * (fp32_error_per_slice(nx) + fp32_error_per_slice(ny));
  return (int)(5e-6 / c);
}
static int default_math(int nx, int ny, int depth, int flags) {
  if (flags & (RSIM_FLAG_MATH_FP32 | RSIM_FLAG_MATH_FP64 | RSIM_FLAG_MATH_COMP32)) return flags;
  return depth > fp32_depth_limit(nx, ny) ? (flags | RSIM_FLAG_MATH_COMP32) : flags;
}
static bool fused_row_len(int n) {
  return n == 256 || n == 512 || n == 1024 || n == 2048 || n == 3072 || n == 4096 || n == 6144 ||
         n == 8192 || n == 16384;
}
static bool fused_col_len(int n) {
  return n == 256 || n == 512 || n == 1024 || n == 2048 || n == 3072 || n == 4096 || n == 6144 ||
         n == 8192 || n == kBigCol;
}

bool fused_supported(int nx, int ny, int dtype) {
  return dtype == RSIM_C64 && fused_row_len(nx) && fused_col_len(ny) &&
         (unsigned long long)nx * ny * 8ull <= 0xFFFFFFFFull;
}

static int propagate_plain(void* u, void* U, int nx, int ny, int dtype, double dx, double dy,
                           double wl, double dz, double freq, bool skip_fft, hipStream_t s) {
  if (!skip_fft) RSIM_CHECK(fft2(dtype, RSIM_FORWARD, u, U, nx, ny, s));
  RSIM_CHECK(chirp_cutoff(U, nx, ny, dtype, dx, dy, wl, dz, freq, true, s));
  RSIM_CHECK(fft2(dtype, RSIM_INVERSE, U, u, nx, ny, s));
  return RSIM_OK;
}

int propagate(void* u, void* U, int nx, int ny, int dtype, double dx, double dy, double wl,
              double dz, double freq, bool skip_fft, hipStream_t s) {
  return propagate_plain(u, U, nx, ny, dtype, dx, dy, wl, dz, freq, skip_fft, s);
}

// History frame k: area_v1 of the field in f at the entry plane of slice sl.
static int hist_frame(const void* f, int nx, int ny, double dx, double dy, double pz, int sl,
                      int k, const HistSpec* h, hipStream_t s, int dtype = RSIM_C64) {
  const double z = h->z_start + sl * pz;   // optical_element.py:548
  return detector_area_v1(f, nx, ny, dtype, dx, dy, h->ox, h->oy, h->dsx, h->dsy, z, 1.0, false,
                          h->frames + (size_t)k * h->oy * h->ox, s);
}

// Tables and transmission arguments of one fused run (all cached; built
// before any launch, and before graph capture in the batched form).
template <typename TW>
struct FusedPrep {
  StageTw<TW> tx{}, ty{};
  const ChirpTables* ct = nullptr;
  TransArgsF ta{};
  int tm = TM_BILINEAR;
  size_t slice_elems = 0;
  const uint8_t* ids0 = nullptr;
  const float2* field0 = nullptr;
  // point the transmission arguments at voxel slice sl
  void at(TransArgsF& a, int sl) const {
    if (tm == TM_FIELD)
      a.dbv = field0 + (size_t)sl * slice_elems;
    else
      a.ids = ids0 + (size_t)sl * slice_elems;
  }
};

static VoxSrc src_of(const rsim_volume* vol) {
  VoxSrc v;
  v.nz = vol->nz;
  v.ny_g = vol->ny_g;
  v.nx_g = vol->nx_g;
  v.vol = vol;
  return v;
}

// Point the transmission arguments at a pair of axis maps; the transmission
// mode they allow (material-id lookup when both maps are integral and affine).
static int use_maps(TransArgsF& ta, const AxisMap& xm, const AxisMap& ym, bool ids) {
  ta.xm = xm.dev;
  ta.ym = ym.dev;
  ta.ax = xm.a;
  ta.cx = xm.c;
  ta.xlo = xm.lo;
  ta.xhi = xm.hi;
  ta.ay = ym.a;
  ta.cy = ym.c;
  ta.ylo = ym.lo;
  ta.yhi = ym.hi;
  if (!ids) return TM_FIELD;
  return xm.all_integral && ym.all_integral && xm.affine && ym.affine ? TM_LUT : TM_BILINEAR;
}

// Tables of a run whose slices all propagate over their thickness pz with an
// unscaled voxel map (cone-beam runs add per-slice tables of their own).
template <typename TW>
static int fused_prepare(int nx, int ny, double dx, double dy, double wl, double freq,
                         const VoxSrc& src, double px, double py, double pz, double xpos,
                         double ypos, FusedPrep<TW>* P) {
  const int tdtype = std::is_same<TW, twc>::value ? kTwComp : sizeof(TW) == 8 ? RSIM_C128 : RSIM_C64;
  const void* twx[4];
  const void* twy[4];
  if (!reg_stage_tables(tdtype, nx, twx) || !reg_stage_tables(tdtype, ny, twy))
    return fail(RSIM_E_HIP, "missing FFT schedule tables");
  using TP = typename std::remove_reference<decltype(P->tx.t[0])>::type;
  for (int i = 0; i < 4; ++i) {
    P->tx.t[i] = reinterpret_cast<TP>(twx[i]);
    P->ty.t[i] = reinterpret_cast<TP>(twy[i]);
  }
  RSIM_CHECK(chirp_tables(nx, ny, dx, dy, wl, pz, freq, &P->ct));
  const AxisMap* xm = nullptr;
  const AxisMap* ym = nullptr;
  RSIM_CHECK(axis_map(nx, dx, 1.0, xpos, src.nx_g, px, &xm));
  RSIM_CHECK(axis_map(ny, dy, 1.0, ypos, src.ny_g, py, &ym));
  const double K = ((2.0 * M_PI) * pz) / wl;
  TransArgsF& ta = P->ta;
  ta.nxg = src.nx_g;
  ta.K = (float)K;
  P->slice_elems = (size_t)src.ny_g * src.nx_g;
  if (src.vol) {
    RSIM_CHECK(material_tables(src.vol, K, &ta.lut, &ta.dbf));
    P->ids0 = src.vol->ids;
    ta.ids = src.vol->ids;
  } else {
    P->field0 = src.field;
    ta.dbv = src.field;
  }
  P->tm = use_maps(ta, *xm, *ym, src.vol != nullptr);
  return RSIM_OK;
}

// Cone-beam runs: each slice has its own propagation distance and voxel-map
// scale, so its own chirp tables and axis maps.  They are built for the whole
// run on the host (in parallel over slices, straight into a pinned staging
// buffer) and uploaded with one copy into one allocation.  A run's set is one
// entry of the bounded table cache, keyed by everything that shapes it, so a
// repeated geometry reuses it and the least recently used sets go first; a
// set above half the cache bound is the run's own, freed on its stream when
// the run ends.  Nothing is cached per slice.
struct SliceTables {
  ChirpTables ct;
  AxisMap xm, ym;
};
struct RunSliceTables {
  std::vector<SliceTables> sets;
  unsigned char* dev = nullptr;
  hipStream_t s = nullptr;
  const void* pinned = nullptr;   // a cached set, held for this run
  ~RunSliceTables() {
    if (dev) (void)hipFreeAsync(dev, s);
    if (pinned) table_pin(pinned, -1);
  }
};

namespace {
// Cached cone-beam table sets: (device, parameter bits..., scales, distances)
// -> the slices' table descriptions pointing into the set's allocation.
struct RunTabEntry {
  std::vector<SliceTables> sets;
  unsigned char* dev = nullptr;
};
std::map<std::vector<uint64_t>, RunTabEntry> g_runtab;
}  // namespace

// Under table_mu (rsim_release_caches): free every cached cone-beam set.
static void release_run_tables() {
  for (auto& kv : g_runtab) (void)hipFree(kv.second.dev);
  g_runtab.clear();
}

namespace {

// Pinned host staging buffer per device, reused once its last upload is done.
struct Staging {
  unsigned char* host = nullptr;
  size_t cap = 0;
  hipEvent_t done = nullptr;
};
std::map<int, Staging> g_staging;
std::mutex g_staging_mu;
}  // namespace

static int build_run_slice_tables(int nx, int ny, double dx, double dy, double wl, double freq,
                                  const VoxSrc& src, double px, double py, double pz,
                                  double xpos, double ypos, int count, const double* scales,
                                  const double* dzs, hipStream_t s, RunSliceTables* rt) {
  const size_t cb = (chirp_bytes(nx, ny) + 255) & ~size_t(255);
  const size_t xb = (axis_map_bytes(nx) + 255) & ~size_t(255);
  const size_t yb = (axis_map_bytes(ny) + 255) & ~size_t(255);
  const size_t per = cb + xb + yb;
  const size_t total = per * (size_t)count;
  std::lock_guard<std::recursive_mutex> tl(table_mu());
  std::vector<uint64_t> key{(uint64_t)current_device(), (uint64_t)nx, (uint64_t)ny, bits(dx), bits(dy),
                            bits(wl), bits(freq), (uint64_t)src.nx_g, (uint64_t)src.ny_g, bits(px),
                            bits(py), bits(pz), bits(xpos), bits(ypos), (uint64_t)count};
  for (int k = 0; k < count; ++k) {
    key.push_back(bits(scales ? scales[k] : 1.0));
    key.push_back(bits(dzs ? dzs[k] : pz));
  }
  auto hit = g_runtab.find(key);
  if (hit != g_runtab.end()) {
    table_touch(hit->second.dev);
    rt->sets = hit->second.sets;
    if (table_pin(hit->second.dev, +1)) rt->pinned = hit->second.dev;
    return RSIM_OK;
  }
  RSIM_CHECK(capture_miss("cone-beam per-slice tables"));
  const bool cache = total <= kTableCacheBytes / 2;
  rt->sets.assign(count, SliceTables{});
  unsigned char* dev = nullptr;
  if (cache) {
    RSIM_HIP(hipMalloc((void**)&dev, total));
  } else {
    RSIM_HIP(hipMallocAsync((void**)&dev, total, s));
    rt->dev = dev;   // the run's own: freed on its stream when the run ends
    rt->s = s;
  }
  std::lock_guard<std::mutex> lk(g_staging_mu);
  Staging& st = g_staging[current_device()];
  if (st.done) RSIM_HIP(hipEventSynchronize(st.done));
  else RSIM_HIP(hipEventCreateWithFlags(&st.done, hipEventDisableTiming));
  if (st.cap < total) {
    if (st.host) RSIM_HIP(hipHostFree(st.host));
    st.host = nullptr;
    st.cap = 0;
    RSIM_HIP(hipHostMalloc((void**)&st.host, total, hipHostMallocDefault));
    st.cap = total;
  }
  const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  const int nthreads = (int)std::min<unsigned>(hw, (unsigned)std::max(1, count / 4));
  auto work = [&](int t) {
    for (int k = t; k < count; k += nthreads) {
      unsigned char* h = st.host + per * (size_t)k;
      const double dz = dzs ? dzs[k] : pz, sc = scales ? scales[k] : 1.0;
      SliceTables& e = rt->sets[k];
      chirp_build(nx, ny, dx, dy, wl, dz, freq, global_phase(wl, dz), h);
      axis_map_build(nx, dx, sc, xpos, src.nx_g, px, &e.xm, h + cb);
      axis_map_build(ny, dy, sc, ypos, src.ny_g, py, &e.ym, h + cb + xb);
      unsigned char* d = dev + per * (size_t)k;
      chirp_bind(&e.ct, d, nx, ny);
      axis_map_bind(&e.xm, d + cb, nx);
      axis_map_bind(&e.ym, d + cb + xb, ny);
    }
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < nthreads; ++t) pool.emplace_back(work, t);
  work(0);
  for (auto& th : pool) th.join();
  RSIM_HIP(hipMemcpyAsync(dev, st.host, total, hipMemcpyHostToDevice, s));
  RSIM_HIP(hipEventRecord(st.done, s));
  if (cache) {
    RSIM_CHECK(table_insert(dev, total, [key] { g_runtab.erase(key); }));
    g_runtab[key] = RunTabEntry{rt->sets, dev};
    if (table_pin(dev, +1)) rt->pinned = dev;
  }
  return RSIM_OK;
}

int release_staging() {
  std::lock_guard<std::mutex> lk(g_staging_mu);
  for (auto& kv : g_staging) {
    if (kv.second.done) {
      (void)hipEventSynchronize(kv.second.done);
      (void)hipEventDestroy(kv.second.done);
    }
    if (kv.second.host) (void)hipHostFree(kv.second.host);
  }
  g_staging.clear();
  return RSIM_OK;
}

template <typename TW>
static int run_fused(void* u, void* U, int nx, int ny, double dx, double dy, double wl,
                     double freq, const VoxSrc& src, double px, double py, double pz,
                     double xpos, double ypos, int first, int count, bool want_spec,
                     const double* scales, const double* dzs, hipStream_t s,
                     const HistSpec* hist) {
  const int total = count;
  const int tdtype = std::is_same<TW, twc>::value ? kTwComp : sizeof(TW) == 8 ? RSIM_C128 : RSIM_C64;
  FusedPrep<TW> P;
  RSIM_CHECK(fused_prepare<TW>(nx, ny, dx, dy, wl, freq, src, px, py, pz, xpos, ypos, &P));
  const StageTw<TW>& tx = P.tx;
  const StageTw<TW>& ty = P.ty;
  const ChirpTables* ct = P.ct;
  TransArgsF ta = P.ta;
  int tm = P.tm;
  // cone-beam runs: the voxel-map scale and the propagation distance change
  // from slice to slice (tables cached per value)
  const bool per_slice = scales || dzs;
  // three-pass columns: 128-point sub-transform tables and the W_ny table
  const bool big = ny == kBigCol;
  BigColTables<TW> bigt;
  if (big) {
    const void* tw128[4];
    if (!reg_stage_tables(tdtype, 128, tw128)) return fail(RSIM_E_HIP, "missing FFT schedule tables");
    using TP = typename std::remove_reference<decltype(bigt.t128.t[0])>::type;
    for (int i = 0; i < 4; ++i) bigt.t128.t[i] = reinterpret_cast<TP>(tw128[i]);
    bigt.wN = reinterpret_cast<const BigTwElem<TW>*>(
        fft_full_table(std::is_same<TW, double>::value ? RSIM_C128 : kTwComp, ny));
    if (!bigt.wN) return fail(RSIM_E_HIP, "missing FFT twiddle table for %d", ny);
  }
  const float spec_scale = (float)((double)nx * (double)ny);
  P.at(ta, first);
  RunSliceTables rts;
  if (per_slice)
    RSIM_CHECK(build_run_slice_tables(nx, ny, dx, dy, wl, freq, src, px, py, pz, xpos, ypos,
                                      count, scales, dzs, s, &rts));
  // History frames after slice 0: the MID pass bins its field in x into xs
  // (nj x ox doubles, stream-ordered for this run); the y stage follows it.
  // Without a separable plan the pass stores the field into u instead.
  XBinArgs xb{};
  bool xbin = false;
  struct XsBuf {
    double* p = nullptr;
    hipStream_t s = nullptr;
    ~XsBuf() { if (p) (void)hipFreeAsync(p, s); }
  } xsb;
  if (hist && count > 1) {
    RSIM_CHECK(capture_miss("history frames of a slice loop"));
    RSIM_CHECK(detector_xbin_plan(nx, ny, dx, dy, hist->ox, hist->oy, hist->dsx, hist->dsy, &xb,
                                  &xbin));
    if (xbin) {
      xsb.s = s;
      RSIM_HIP(hipMallocAsync((void**)&xsb.p, (size_t)xb.nj * xb.ox * sizeof(double), s));
      xb.xs = xsb.p;
    }
  }
  // Intermediate layout.  In place: every pass reads and writes U.  Tiled
  // (where the tuner found it faster): the column pass stores its output as
  // column tiles into u, which the next row pass reads back into U; the last
  // column pass stores row-major.  Not with the three-pass columns, nor when
  // the MID pass stores history fields into u, nor when the launch shapes
  // may change from slice to slice (cone-beam tables).
  const bool may_tile = !big && !per_slice && !(hist && !xbin);
  int rv = 0, cv = 0, tile = 0;
  RSIM_CHECK(fused_schedule<TW>(nx, ny, tm, tx, ty, ta, ct, s, &rv, &cv, may_tile ? &tile : nullptr,
                                big ? &bigt : nullptr));
  int w_in = 0;   // tile width of the W that the next row pass reads (0: W in U)
  for (int k = 0; k < count; ++k) {
    const int sl = first + k;
    if (cancel_requested()) {
      // stop at this slice boundary: the row IFFT of W of slice k - 1 leaves
      // u = the field after k slices (U then does not hold FFT2(u))
      if (k > 0) {
        if (w_in) {
          fused_row(nx, rv, ROW_LAST, tm, u, U, ny, tx, ta, s, AdjArgs{}, w_in);
          RSIM_HIP(hipMemcpyAsync(u, U, (size_t)nx * ny * 8, hipMemcpyDeviceToDevice, s));
        } else {
          fused_row(nx, rv, ROW_LAST, tm, U, u, ny, tx, ta, s);
        }
        RSIM_LAUNCHED();
      }
      note_done(k, total);
      return fail(RSIM_E_CANCELLED, "multislice cancelled after %d of %d slices (at slice %d)",
                  k, total, sl);
    }
    if (per_slice) {
      const SliceTables& e = rts.sets[k];
      ct = &e.ct;
      tm = use_maps(ta, e.xm, e.ym, src.vol != nullptr);
      RSIM_CHECK(fused_schedule<TW>(nx, ny, tm, tx, ty, ta, ct, s, &rv, &cv,
                                    may_tile ? &tile : nullptr, big ? &bigt : nullptr));
    }
    P.at(ta, sl);
    const bool last = k == count - 1;
    // History frame of the field entering this slice: the input for slice 0;
    // otherwise the row IFFT of W that the MID pass holds in registers,
    // binned in x by that pass (or stored into u, which the loop does not
    // read after slice 0)
    if (hist && k == 0) RSIM_CHECK(hist_frame(u, nx, ny, dx, dy, pz, sl, k, hist, s));
    AdjArgs store{};
    if (hist && k > 0) {
      if (xbin) store.xb = xb;
      else store.fld_out = reinterpret_cast<cpx<float>*>(u);
    }
    // the row-pass samples are MID launches (FIRST and LAST run one transform)
    if (k > 0) ktime_begin(KT_ROW, s);
    const void* rin = k == 0 ? u : (w_in ? u : U);
    if (!fused_row(nx, rv, k == 0 ? ROW_FIRST : ROW_MID, tm, rin, U, ny, tx, ta, s, store,
                   k == 0 ? 0 : w_in))
      return fail(RSIM_E_HIP, "no fused row kernel for nx=%d", nx);
    ktime_end(KT_ROW, s);
    RSIM_LAUNCHED();
    if (hist && k > 0) {
      if (xbin)
        RSIM_CHECK(detector_ybin(xb.xs, nx, ny, dx, dy, hist->ox, hist->oy, hist->dsx, hist->dsy,
                                 hist->z_start + sl * pz, 1.0, false,
                                 hist->frames + (size_t)k * hist->oy * hist->ox, s));
      else
        RSIM_CHECK(hist_frame(u, nx, ny, dx, dy, pz, sl, k, hist, s));
    }
    ktime_begin(KT_COL, s);
    if (big) {
      // in place in U; the spectrum of the final field is formed after the loop
      if (!big_col_go<TW>(cv, U, nx, ny, bigt.t128, bigt.wN, ct->Bx, ct->Ay, ct->Rmax, s))
        return fail(RSIM_E_HIP, "no three-pass column schedule for ny=%d", ny);
    } else {
      // last slice: W -> u (and the spectrum -> U); otherwise in place in U,
      // or as tiles into u
      const int w_out = last ? 0 : tile;
      if (!fused_col(ny, cv, last && want_spec, U, (last || w_out) ? u : U, U, nx, ty, ct->Bx,
                     ct->Ay, ct->Rmax, spec_scale, s, w_out))
        return fail(RSIM_E_HIP, "no fused column kernel for ny=%d", ny);
      w_in = w_out;
    }
    ktime_end(KT_COL, s);
    RSIM_LAUNCHED();
    if (last) {
      // row IFFT of W
      fused_row(nx, rv, ROW_LAST, tm, big ? U : u, u, ny, tx, ta, s);
      RSIM_LAUNCHED();
      if (big && want_spec) RSIM_CHECK(fft2(RSIM_C64, RSIM_FORWARD, u, U, nx, ny, s));
    }
    if (progress_due(k + 1, total)) {
      // the field after slice k is reported only once it exists: for a
      // slice other than the last it is W in U, still one row IFFT away,
      // which a cancel at the next slice boundary performs
      RSIM_HIP(hipStreamSynchronize(s));
      report_progress("sample_slice", k + 1, total);
    }
  }
  note_done(count, total);
  return RSIM_OK;
}

int slices_apply(void* u, void* U, int nx, int ny, int dtype, double dx, double dy, double wl,
                 double freq, const VoxSrc& src, double px, double py, double pz, double xpos,
                 double ypos, int first, int count, const double* scales, const double* dzs,
                 int engine, int flags, hipStream_t s, const HistSpec* hist) {
  const int total = count;
  if (count == 0) return RSIM_OK;
  const bool want_spec = !(flags & RSIM_FLAG_NO_SPECTRUM);
  if (engine == RSIM_PLAIN || !fused_supported(nx, ny, dtype)) {
    for (int k = 0; k < count; ++k) {
      if (cancel_requested()) {
        note_done(k, total);
        return fail(RSIM_E_CANCELLED, "multislice cancelled after %d of %d slices (at slice %d)",
                    k, total, first + k);
      }
      if (hist) RSIM_CHECK(hist_frame(u, nx, ny, dx, dy, pz, first + k, k, hist, s, dtype));
      RSIM_CHECK(src.plain(u, dtype, first + k, scales ? scales[k] : 1.0, s));
      RSIM_CHECK(propagate_plain(u, U, nx, ny, dtype, dx, dy, wl, dzs ? dzs[k] : pz, freq, false, s));
      if (progress_due(k + 1, total)) {
        RSIM_HIP(hipStreamSynchronize(s));
        report_progress("sample_slice", k + 1, total);
      }
    }
    note_done(count, total);
    return RSIM_OK;
  }
  // ---- fused complex64 engine ----
  // Butterfly arithmetic: FP32 while its error over the volume's depth is
  // predicted within half the 1e-5 bar, compensated FP32 beyond.
  flags = default_math(nx, ny, src.nz, flags);
  if (flags & RSIM_FLAG_MATH_COMP32)
    return run_fused<twc>(u, U, nx, ny, dx, dy, wl, freq, src, px, py, pz, xpos, ypos, first,
                          count, want_spec, scales, dzs, s, hist);
  if (flags & RSIM_FLAG_MATH_FP64)
    return run_fused<double>(u, U, nx, ny, dx, dy, wl, freq, src, px, py, pz, xpos, ypos, first,
                             count, want_spec, scales, dzs, s, hist);
  return run_fused<float>(u, U, nx, ny, dx, dy, wl, freq, src, px, py, pz, xpos, ypos, first,
                          count, want_spec, scales, dzs, s, hist);
}

// ---------------------------------------------------------------------------
// Propagation chains (rsim_chain_apply_2d): an optional band limit, then
// sample steps and free-space steps in order.  On the fused engine the whole
// chain is one sequence of row and column passes: a propagation with no
// transmission after it is not completed in real space but carried into the
// next column pass, whose distance is the sum, P(a) P(b) = P(a + b) (chirps
// multiply, the cutoff mask is idempotent; big-wave/propagation.py:396-442).
// ---------------------------------------------------------------------------
namespace {
// One pass of a fused chain: a row pass (FIRST, MID or LAST) with the
// transmission of slice k of sample step j (j < 0: none), or a column pass
// whose chirp tables (one per instance: each has its own wavelength,
// distances and cutoff) cover the propagations folded into it.
struct ChainOp {
  bool col;
  int mode, j, k;
  std::vector<const ChirpTables*> ct;
};
double2 cmul2(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
}  // namespace

// The chain of every instance (all of one grid, with the same step kinds and
// slice counts) as one pass sequence; each pass launches the instances
// together, up to kMaxBatch per launch (and one launch per transmission mode
// where their voxel maps differ in it).
template <typename TW>
static int run_chain(const rsim_chain_instance* in, int count, int nx, int ny, bool band_limit,
                     int n_steps, bool want_spec, hipStream_t s) {
  const rsim_chain_step* st0 = in[0].steps;
  // Every cached table the passes will read is pinned as soon as it is made,
  // so that making the later ones cannot evict it before the launches.
  struct Pins {
    std::vector<const void*> p;
    void add(const void* x) {
      if (x && table_pin(x, +1)) p.push_back(x);
    }
    ~Pins() {
      for (const void* x : p) table_pin(x, -1);
    }
  } pins;
  // per instance and sample step: tables, transmission arguments, launch shapes
  std::vector<std::vector<FusedPrep<TW>>> prep(count, std::vector<FusedPrep<TW>>(n_steps));
  std::vector<std::vector<int>> rvs(count, std::vector<int>(n_steps, 0));
  int rv0 = 0, cv0 = 0;
  bool have_ref = false;   // the first sample step's shapes serve the passes without transmission
  for (int i = 0; i < count; ++i)
    for (int j = 0; j < n_steps; ++j) {
      const rsim_chain_step& st = in[i].steps[j];
      if (!st.vol || st.n_slices == 0) continue;
      FusedPrep<TW>& P = prep[i][j];
      {
        std::lock_guard<std::recursive_mutex> tl(table_mu());
        RSIM_CHECK(fused_prepare<TW>(nx, ny, in[i].dx, in[i].dy, in[i].wl, in[i].cutoff_freq,
                                     src_of(st.vol), st.pixel_size_x, st.pixel_size_y,
                                     st.pixel_size_z, st.x_pos, st.y_pos, &P));
        pins.add(P.ct->Bx);
        pins.add(P.ta.xm.frac);
        pins.add(P.ta.ym.frac);
      }
      P.at(P.ta, st.first_slice);
      int cv = 0;
      RSIM_CHECK(fused_schedule<TW>(nx, ny, P.tm, P.tx, P.ty, P.ta, P.ct, s, &rvs[i][j], &cv));
      if (!have_ref) {
        rv0 = rvs[i][j];
        cv0 = cv;
        have_ref = true;
      }
    }
  // the pass sequence (the same for every instance) and each instance's
  // folded distance and global phase per column pass
  std::vector<ChainOp> ops;
  std::vector<double> dist(count, 0.0);
  std::vector<double2> gp(count, make_double2(1.0, 0.0));
  bool pending = false;   // a propagation waits for its column pass
  auto row = [&](int mode, int j, int k) { ops.push_back(ChainOp{false, mode, j, k, {}}); };
  auto col = [&]() -> int {
    ChainOp op{true, 0, -1, 0, std::vector<const ChirpTables*>(count, nullptr)};
    std::lock_guard<std::recursive_mutex> tl(table_mu());
    for (int i = 0; i < count; ++i) {
      RSIM_CHECK(chirp_tables(nx, ny, in[i].dx, in[i].dy, in[i].wl, dist[i], in[i].cutoff_freq,
                              &op.ct[i], &gp[i]));
      pins.add(op.ct[i]->Bx);
    }
    ops.push_back(std::move(op));
    return RSIM_OK;
  };
  if (band_limit) {
    row(ROW_FIRST, -1, 0);
    pending = true;
    for (int i = 0; i < count; ++i) {
      dist[i] = 0.0;
      gp[i] = make_double2(1.0, 0.0);
    }
  }
  for (int j = 0; j < n_steps; ++j) {
    if (!st0[j].vol) {
      if (!pending) row(ROW_FIRST, -1, 0);
      for (int i = 0; i < count; ++i) {
        const double dz = in[i].steps[j].dz;
        const double2 g = global_phase(in[i].wl, dz);
        dist[i] = pending ? dist[i] + dz : dz;
        gp[i] = pending ? cmul2(gp[i], g) : g;
      }
      pending = true;
      continue;
    }
    for (int k = 0; k < st0[j].n_slices; ++k) {
      if (pending) {
        RSIM_CHECK(col());
        row(ROW_MID, j, k);
      } else {
        row(ROW_FIRST, j, k);
      }
      pending = true;
      for (int i = 0; i < count; ++i) {
        dist[i] = in[i].steps[j].pixel_size_z;
        gp[i] = global_phase(in[i].wl, dist[i]);
      }
    }
  }
  if (pending) {
    RSIM_CHECK(col());
    row(ROW_LAST, -1, 0);
  }
  // twiddle tables (shared: one grid, one arithmetic)
  StageTw<TW> tx{}, ty{};
  {
    const int tdtype = std::is_same<TW, twc>::value ? kTwComp : sizeof(TW) == 8 ? RSIM_C128 : RSIM_C64;
    const void* twx[4];
    const void* twy[4];
    if (!reg_stage_tables(tdtype, nx, twx) || !reg_stage_tables(tdtype, ny, twy))
      return fail(RSIM_E_HIP, "missing FFT schedule tables");
    using TP = typename std::remove_reference<decltype(tx.t[0])>::type;
    for (int q = 0; q < 4; ++q) {
      tx.t[q] = reinterpret_cast<TP>(twx[q]);
      ty.t[q] = reinterpret_cast<TP>(twy[q]);
    }
  }
  const float spec_scale = (float)((double)nx * (double)ny);
  // field flow per instance: FIRST u -> U; columns and MID in place in U; the
  // last column pass W -> u (spectrum -> U); LAST u -> u
  for (size_t o = 0; o < ops.size(); ++o) {
    const ChainOp& op = ops[o];
    if (op.col) {
      const bool last = o + 1 == ops.size() - 1;
      for (int c0 = 0; c0 < count; c0 += kMaxBatch) {
        const int cnt = std::min(kMaxBatch, count - c0);
        ColBatch b{};
        for (int q = 0; q < cnt; ++q) {
          const int i = c0 + q;
          const ChirpTables* ct = op.ct[i];
          b.inst[q] = ColInst{reinterpret_cast<const cpx<float>*>(in[i].U),
                              reinterpret_cast<cpx<float>*>(last ? in[i].u : in[i].U),
                              reinterpret_cast<cpx<float>*>(in[i].U), ct->Bx, ct->Ay, ct->Rmax};
        }
        ktime_begin(KT_COL, s);
        const bool ok = fused_col_b<TW>(ny, cv0, last && want_spec, b, cnt, nx, ty, spec_scale, s);
        ktime_end(KT_COL, s);
        if (!ok) return fail(RSIM_E_HIP, "no fused column kernel for ny=%d", ny);
        RSIM_LAUNCHED();
      }
      continue;
    }
    // row pass: the instances of each transmission mode together
    int modes[8];
    int nmodes = 0;
    auto tm_of = [&](int i) { return op.j >= 0 ? prep[i][op.j].tm : op.mode == ROW_LAST ? TM_LUT : TM_NONE; };
    for (int i = 0; i < count; ++i) {
      const int tm = tm_of(i);
      bool seen = false;
      for (int q = 0; q < nmodes; ++q) seen = seen || modes[q] == tm;
      if (!seen) modes[nmodes++] = tm;
    }
    for (int q = 0; q < nmodes; ++q) {
      const int tm = modes[q];
      std::vector<int> idx;
      for (int i = 0; i < count; ++i)
        if (tm_of(i) == tm) idx.push_back(i);
      const int rv = op.j >= 0 ? rvs[idx[0]][op.j] : rv0;
      for (size_t c0 = 0; c0 < idx.size(); c0 += kMaxBatch) {
        const int cnt = (int)std::min<size_t>(kMaxBatch, idx.size() - c0);
        RowBatch b{};
        for (int r = 0; r < cnt; ++r) {
          const int i = idx[c0 + r];
          const void* src = op.mode == ROW_MID ? in[i].U : in[i].u;
          void* dst = op.mode == ROW_LAST ? in[i].u : in[i].U;
          b.inst[r].in = reinterpret_cast<const cpx<float>*>(src);
          b.inst[r].out = reinterpret_cast<cpx<float>*>(dst);
          if (op.j >= 0) {
            const FusedPrep<TW>& P = prep[i][op.j];
            b.inst[r].ta = P.ta;
            P.at(b.inst[r].ta, in[i].steps[op.j].first_slice + op.k);
          }
        }
        if (op.mode == ROW_MID) ktime_begin(KT_ROW, s);
        const bool ok = fused_row_b<TW>(nx, rv, op.mode, tm, b, cnt, ny, tx, s);
        ktime_end(KT_ROW, s);
        if (!ok) return fail(RSIM_E_HIP, "no fused row kernel for nx=%d", nx);
        RSIM_LAUNCHED();
      }
    }
  }
  if (!pending && want_spec)
    for (int i = 0; i < count; ++i) RSIM_CHECK(fft2(RSIM_C64, RSIM_FORWARD, in[i].u, in[i].U, nx, ny, s));
  return RSIM_OK;
}

int chain_apply(const rsim_chain_instance* in, int count, int nx, int ny, int dtype,
                bool band_limit, int n_steps, int engine, int flags, hipStream_t s) {
  if (cancel_requested()) return fail(RSIM_E_CANCELLED, "propagation chain cancelled before it started");
  if (count == 0) return RSIM_OK;
  const bool want_spec = !(flags & RSIM_FLAG_NO_SPECTRUM);
  int depth = 0;
  for (int j = 0; j < n_steps; ++j)
    if (in[0].steps[j].vol) depth += in[0].steps[j].n_slices;
  if (engine == RSIM_FUSED && fused_supported(nx, ny, dtype) && ny != kBigCol) {
    flags = default_math(nx, ny, depth, flags);
    if (flags & RSIM_FLAG_MATH_COMP32)
      return run_chain<twc>(in, count, nx, ny, band_limit, n_steps, want_spec, s);
    if (flags & RSIM_FLAG_MATH_FP64)
      return run_chain<double>(in, count, nx, ny, band_limit, n_steps, want_spec, s);
    return run_chain<float>(in, count, nx, ny, band_limit, n_steps, want_spec, s);
  }
  // step by step: the source's band limit leaves U undefined
  // (big-wave/source.py:38-40); every later step leaves U = FFT2(u)
  for (int i = 0; i < count; ++i) {
    const rsim_chain_instance& q = in[i];
    bool spectrum_valid = false;
    if (band_limit) {
      RSIM_CHECK(fft2(dtype, RSIM_FORWARD, q.u, q.U, nx, ny, s));
      RSIM_CHECK(chirp_cutoff(q.U, nx, ny, dtype, q.dx, q.dy, 1.0, 0.0, q.cutoff_freq, false, s));
      RSIM_CHECK(fft2(dtype, RSIM_INVERSE, q.U, q.u, nx, ny, s));
    }
    for (int j = 0; j < n_steps; ++j) {
      const rsim_chain_step& st = q.steps[j];
      if (st.vol) {
        RSIM_CHECK(sample_apply(q.u, q.U, nx, ny, dtype, q.dx, q.dy, q.wl, q.cutoff_freq, st.vol,
                                st.pixel_size_x, st.pixel_size_y, st.pixel_size_z, st.x_pos,
                                st.y_pos, st.first_slice, st.n_slices, nullptr, nullptr, engine,
                                flags & ~RSIM_FLAG_NO_SPECTRUM, s, nullptr));
      } else {
        RSIM_CHECK(propagate(q.u, q.U, nx, ny, dtype, q.dx, q.dy, q.wl, st.dz, q.cutoff_freq,
                             spectrum_valid, s));
      }
      if (!st.vol || st.n_slices > 0) spectrum_valid = true;   // a step that ran leaves FFT2(u)
    }
    if (want_spec && !spectrum_valid) RSIM_CHECK(fft2(dtype, RSIM_FORWARD, q.u, q.U, nx, ny, s));
  }
  return RSIM_OK;
}

// ---------------------------------------------------------------------------
// Differentiable path on the fused engine: per-voxel delta/beta planes on the
// field grid (TM_PLANAR), single-pass column lengths, butterfly arithmetic as
// in the voxel engine (FP32, compensated FP32 or FP64 twiddles).
// ---------------------------------------------------------------------------
bool fused_db_supported(int nx, int ny, int dtype) {
  return fused_supported(nx, ny, dtype) && ny != kBigCol;
}

namespace {
template <typename TW>
struct DbPrep {
  StageTw<TW> tx{}, ty{};
  const ChirpTables* fwd = nullptr;   // chirp over +pz
  const ChirpTables* adj = nullptr;   // chirp over -pz (conjugate, same mask)
  TransArgsF ta{};
  int rv = 0, cv = 0;
  size_t vs = 0;
  const float* d0 = nullptr;
  const float* b0 = nullptr;
  void at(int sl) {
    ta.dpl = d0 + (size_t)sl * vs;
    ta.bpl = b0 + (size_t)sl * vs;
  }
};

template <typename TW>
int db_prepare(int nx, int ny, double dx, double dy, double wl, double freq, const float* delta,
               const float* beta, double pz, bool adjoint, DbPrep<TW>* P, hipStream_t s) {
  const int tdtype = std::is_same<TW, twc>::value ? kTwComp : sizeof(TW) == 8 ? RSIM_C128 : RSIM_C64;
  const void* twx[4];
  const void* twy[4];
  if (!reg_stage_tables(tdtype, nx, twx) || !reg_stage_tables(tdtype, ny, twy))
    return fail(RSIM_E_HIP, "missing FFT schedule tables");
  using TP = typename std::remove_reference<decltype(P->tx.t[0])>::type;
  for (int i = 0; i < 4; ++i) {
    P->tx.t[i] = reinterpret_cast<TP>(twx[i]);
    P->ty.t[i] = reinterpret_cast<TP>(twy[i]);
  }
  RSIM_CHECK(chirp_tables(nx, ny, dx, dy, wl, pz, freq, &P->fwd));
  if (adjoint) RSIM_CHECK(chirp_tables(nx, ny, dx, dy, wl, -pz, freq, &P->adj));
  P->vs = (size_t)nx * ny;
  P->d0 = delta;
  P->b0 = beta;
  P->ta.nxg = nx;
  P->ta.K = (float)(((2.0 * M_PI) * pz) / wl);
  P->at(0);
  return fused_schedule<TW>(nx, ny, TM_PLANAR, P->tx, P->ty, P->ta, P->fwd, s, &P->rv, &P->cv);
}

template <typename TW>
bool db_row(const DbPrep<TW>& P, int nx, int ny, int mode, int tm, const void* in, void* out,
            hipStream_t s, const AdjArgs& aj = AdjArgs{}) {
  ktime_begin(KT_ROW, s);
  const bool ok = fused_row<TW>(nx, P.rv, mode, tm, in, out, ny, P.tx, P.ta, s, aj);
  ktime_end(KT_ROW, s);
  return ok;
}

template <typename TW>
bool db_col(const DbPrep<TW>& P, int nx, int ny, const ChirpTables* ct, const void* in, void* out,
            void* spec, hipStream_t s) {
  ktime_begin(KT_COL, s);
  const bool ok = fused_col<TW>(ny, P.cv, spec != nullptr, in, out, spec, nx, P.ty, ct->Bx,
                                   ct->Ay, ct->Rmax, (float)((double)nx * (double)ny), s);
  ktime_end(KT_COL, s);
  return ok;
}
// Forward: the fused slice chain FIRST, col, (MID, col)..., LAST.  The field
// entering slice k (k % every == 0) is stored by that slice's MID pass (the
// input itself for k = 0); the last column pass also leaves FFT2(u) in U.
template <typename TW>
int db_forward(void* u, void* U, int nx, int ny, double dx, double dy, double wl, double freq,
               const float* delta, const float* beta, double pz, int n, void* ckpt, int every,
               hipStream_t s) {
  DbPrep<TW> P;
  RSIM_CHECK(db_prepare(nx, ny, dx, dy, wl, freq, delta, beta, pz, false, &P, s));
  const size_t fb = (size_t)nx * ny * 8;
  if (ckpt) RSIM_HIP(hipMemcpyAsync(ckpt, u, fb, hipMemcpyDeviceToDevice, s));
  for (int k = 0; k < n; ++k) {
    if (cancel_requested()) return fail(RSIM_E_CANCELLED, "multislice cancelled at slice %d", k);
    P.at(k);
    AdjArgs aj{};
    if (ckpt && k > 0 && k % every == 0)
      aj.fld_out = reinterpret_cast<cpx<float>*>((char*)ckpt + (size_t)(k / every) * fb);
    const bool last = k == n - 1;
    if (!db_row(P, nx, ny, k == 0 ? ROW_FIRST : ROW_MID, TM_PLANAR, k == 0 ? u : U, U, s, aj))
      return fail(RSIM_E_HIP, "no fused row kernel for nx=%d", nx);
    RSIM_LAUNCHED();
    if (!db_col(P, nx, ny, P.fwd, U, last ? u : U, last ? U : nullptr, s))
      return fail(RSIM_E_HIP, "no fused column kernel for ny=%d", ny);
    RSIM_LAUNCHED();
    if (progress_due(k + 1, n)) {
      RSIM_HIP(hipStreamSynchronize(s));
      report_progress("multislice", k + 1, n);
    }
  }
  if (!db_row(P, nx, ny, ROW_LAST, TM_PLANAR, u, u, s))
    return fail(RSIM_E_HIP, "no fused row kernel for nx=%d", nx);
  RSIM_LAUNCHED();
  return RSIM_OK;
}

// Adjoint: one chain over all slices in the spectral row layout,
//   V = FFT_x(g);  per slice s from the last: W = IFFT_y(conj(M) FFT_y(V)),
//   then ADJ (or ADJ_LAST for s = 0) with the stored field u_s,
// with the fields of each checkpoint segment rebuilt by a forward chain of
// their own in a set of work slots (the last slot doubles as that chain's
// buffer).  The rebuilds run on a second stream, one segment ahead of the
// adjoint chain, in two alternating slot sets: events order a rebuild after
// the adjoint of the segment that last used its set, and the adjoint of a
// segment after its rebuild.
template <typename TW>
int db_adjoint(void* g, void* U, int nx, int ny, double dx, double dy, double wl, double freq,
               const float* delta, const float* beta, double pz, int n, const void* ckpt,
               int every, void* work, float* gdelta, float* gbeta, hipStream_t s) {
  DbPrep<TW> P;
  RSIM_CHECK(db_prepare(nx, ny, dx, dy, wl, freq, delta, beta, pz, true, &P, s));
  const size_t fb = (size_t)nx * ny * 8;
  const int nseg = (n + every - 1) / every;
  const size_t set_b = (size_t)(std::min(every, n) - 1) * fb;
  hipStream_t s2 = side_stream();
  if (!s2) return fail(RSIM_E_HIP, "could not create the rebuild stream");
  // events: [0] the caller's prior work, then rebuilt[seg], consumed[seg]
  std::vector<hipEvent_t> ev(1 + 2 * (size_t)nseg, nullptr);
  auto cleanup = [&]() {
    for (hipEvent_t e : ev)
      if (e) hipEventDestroy(e);
  };
  for (hipEvent_t& e : ev) {
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
      cleanup();
      return fail(RSIM_E_HIP, "hipEventCreateWithFlags failed");
    }
  }
  hipEvent_t* rebuilt = ev.data() + 1;
  hipEvent_t* consumed = ev.data() + 1 + nseg;
  auto slots = [&](int seg) { return (char*)work + (size_t)((nseg - 1 - seg) & 1) * set_b; };
  auto field = [&](int seg, int k) -> void* {
    const int s0 = seg * every;
    return k == s0 ? (void*)((const char*)ckpt + (size_t)seg * fb)
                   : (void*)(slots(seg) + (size_t)(k - s0 - 1) * fb);
  };
  // rebuild the fields s0+1 .. s1-1 of segment seg on s2
  auto rebuild = [&](int seg) -> int {
    const int s0 = seg * every;
    const int s1 = std::min(n, s0 + every);
    if (seg + 2 < nseg) RSIM_HIP(hipStreamWaitEvent(s2, consumed[seg + 2], 0));
    if (s1 - s0 >= 2) {
      void* R = field(seg, s1 - 1);
      const void* ck = field(seg, s0);
      P.at(s0);
      bool ok = db_row(P, nx, ny, ROW_FIRST, TM_PLANAR, ck, R, s2) &&
                db_col(P, nx, ny, P.fwd, R, R, nullptr, s2);
      for (int k = s0 + 1; ok && k < s1 - 1; ++k) {
        P.at(k);
        AdjArgs aj{};
        aj.fld_out = reinterpret_cast<cpx<float>*>(field(seg, k));
        ok = db_row(P, nx, ny, ROW_MID, TM_PLANAR, R, R, s2, aj) &&
             db_col(P, nx, ny, P.fwd, R, R, nullptr, s2);
      }
      ok = ok && db_row(P, nx, ny, ROW_LAST, TM_PLANAR, R, R, s2);
      if (!ok) return fail(RSIM_E_HIP, "no fused kernels for %dx%d", nx, ny);
      RSIM_LAUNCHED();
    }
    RSIM_HIP(hipEventRecord(rebuilt[seg], s2));
    return RSIM_OK;
  };
  auto body = [&]() -> int {
    // the rebuilds read checkpoints written by work already on s
    RSIM_HIP(hipEventRecord(ev[0], s));
    RSIM_HIP(hipStreamWaitEvent(s2, ev[0], 0));
    if (!db_row(P, nx, ny, ROW_FIRST, TM_NONE, g, U, s))
      return fail(RSIM_E_HIP, "no fused row kernel for nx=%d", nx);
    RSIM_LAUNCHED();
    RSIM_CHECK(rebuild(nseg - 1));
    long long done = 0;
    for (int seg = nseg - 1; seg >= 0; --seg) {
      if (seg > 0) RSIM_CHECK(rebuild(seg - 1));
      RSIM_HIP(hipStreamWaitEvent(s, rebuilt[seg], 0));
      const int s0 = seg * every;
      const int s1 = std::min(n, s0 + every);
      for (int k = s1 - 1; k >= s0; --k) {
        if (cancel_requested()) return fail(RSIM_E_CANCELLED, "adjoint cancelled at slice %d", k);
        if (!db_col(P, nx, ny, P.adj, U, U, nullptr, s))
          return fail(RSIM_E_HIP, "no fused column kernel for ny=%d", ny);
        P.at(k);
        AdjArgs aj{};
        aj.fld_in = reinterpret_cast<const cpx<float>*>(field(seg, k));
        aj.gd = gdelta + (size_t)k * P.vs;
        aj.gb = gbeta + (size_t)k * P.vs;
        if (!db_row(P, nx, ny, k == 0 ? ROW_ADJ_LAST : ROW_ADJ, TM_PLANAR, U, k == 0 ? g : U, s,
                    aj))
          return fail(RSIM_E_HIP, "no fused adjoint row kernel for nx=%d", nx);
        RSIM_LAUNCHED();
        ++done;
        if (progress_due(done, n)) {
          RSIM_HIP(hipStreamSynchronize(s));
          report_progress("multislice_adjoint", done, n);
        }
      }
      RSIM_HIP(hipEventRecord(consumed[seg], s));
    }
    return RSIM_OK;
  };
  int rc = body();
  if (rc != RSIM_OK) {
    // leave no rebuild running into buffers the caller may free
    hipStreamSynchronize(s2);
    hipStreamSynchronize(s);
  }
  cleanup();
  return rc;
}
}  // namespace

int fused_multislice_db(void* u, void* U, int nx, int ny, double dx, double dy, double wl,
                        double freq, const float* delta, const float* beta, double pz, int n,
                        void* ckpt, int every, int flags, hipStream_t s) {
  if (flags & RSIM_FLAG_MATH_COMP32)
    return db_forward<twc>(u, U, nx, ny, dx, dy, wl, freq, delta, beta, pz, n, ckpt, every, s);
  if (flags & RSIM_FLAG_MATH_FP64)
    return db_forward<double>(u, U, nx, ny, dx, dy, wl, freq, delta, beta, pz, n, ckpt, every, s);
  return db_forward<float>(u, U, nx, ny, dx, dy, wl, freq, delta, beta, pz, n, ckpt, every, s);
}

int fused_multislice_db_adjoint(void* g, void* U, int nx, int ny, double dx, double dy,
                                double wl, double freq, const float* delta, const float* beta,
                                double pz, int n, const void* ckpt, int every, void* work,
                                float* gdelta, float* gbeta, int flags, hipStream_t s) {
  if (flags & RSIM_FLAG_MATH_COMP32)
    return db_adjoint<twc>(g, U, nx, ny, dx, dy, wl, freq, delta, beta, pz, n, ckpt, every, work,
                           gdelta, gbeta, s);
  if (flags & RSIM_FLAG_MATH_FP64)
    return db_adjoint<double>(g, U, nx, ny, dx, dy, wl, freq, delta, beta, pz, n, ckpt, every,
                              work, gdelta, gbeta, s);
  return db_adjoint<float>(g, U, nx, ny, dx, dy, wl, freq, delta, beta, pz, n, ckpt, every, work,
                           gdelta, gbeta, s);
}

int sample_apply(void* u, void* U, int nx, int ny, int dtype, double dx, double dy, double wl,
                 double freq, const rsim_volume* vol, double px, double py, double pz,
                 double xpos, double ypos, int first, int count, const double* scales,
                 const double* dzs, int engine, int flags, hipStream_t s, const HistSpec* hist) {
  VoxSrc src = src_of(vol);
  src.plain = [&](void* f, int dt, int slice, double scale, hipStream_t st) {
    return transmission(f, nx, ny, dt, dx, dy, wl, vol, slice, px, py, pz, xpos, ypos, scale, st);
  };
  return slices_apply(u, U, nx, ny, dtype, dx, dy, wl, freq, src, px, py, pz, xpos, ypos, first,
                      count, scales, dzs, engine, flags, s, hist);
}

// ---------------------------------------------------------------------------
// Batched fused multislice: instances grouped by shape into launches of up to
// kMaxBatch fields, the whole slice loop captured as one hipGraph.
// ---------------------------------------------------------------------------
template <typename TW>
static int batch_capture(const rsim_batch_instance* in, int count, int first, int n,
                         bool want_spec, hipStream_t cs, hipGraph_t* graph, int* ngroups,
                         int* nkern) {
  std::vector<FusedPrep<TW>> prep(count);
  for (int i = 0; i < count; ++i) {
    const rsim_batch_instance& q = in[i];
    if (!fused_supported(q.nx, q.ny, RSIM_C64) || q.ny == kBigCol)
      return fail(RSIM_E_UNSUPPORTED, "instance %d: %dx%d has no batched fused schedule", i, q.nx,
                  q.ny);
    RSIM_CHECK(fused_prepare<TW>(q.nx, q.ny, q.dx, q.dy, q.wl, q.cutoff_freq, src_of(q.vol),
                                 q.pixel_size_x, q.pixel_size_y, q.pixel_size_z, q.x_pos,
                                 q.y_pos, &prep[i]));
  }
  struct Group {
    int nx, ny;
    int tm;
    int rv, cv;
    std::vector<int> idx;
  };
  std::vector<Group> groups;
  for (int i = 0; i < count; ++i) {
    int gi = -1;
    for (size_t g = 0; g < groups.size(); ++g)
      if (groups[g].nx == in[i].nx && groups[g].ny == in[i].ny && groups[g].tm == prep[i].tm)
        gi = (int)g;
    if (gi < 0) {
      groups.push_back(Group{in[i].nx, in[i].ny, prep[i].tm, 0, 0, {}});
      gi = (int)groups.size() - 1;
    }
    groups[gi].idx.push_back(i);
  }
  // launch shapes: autotuned (scratch allocation and timing) before capture
  for (auto& g : groups) {
    const FusedPrep<TW>& P = prep[g.idx[0]];
    RSIM_CHECK(fused_schedule<TW>(g.nx, g.ny, g.tm, P.tx, P.ty, P.ta, P.ct, cs, &g.rv, &g.cv));
  }
  RSIM_HIP(hipStreamSynchronize(cs));

  ktime_suspend(true);
  hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) {
    ktime_suspend(false);
    return fail(RSIM_E_HIP, "hipStreamBeginCapture failed: %s", hipGetErrorString(e));
  }
  int kern = 0;
  bool ok = true;
  auto rows = [&](int k, int mode) {
    for (auto& g : groups)
      for (size_t c0 = 0; c0 < g.idx.size() && ok; c0 += kMaxBatch) {
        const int cnt = (int)std::min<size_t>(kMaxBatch, g.idx.size() - c0);
        RowBatch b{};
        for (int j = 0; j < cnt; ++j) {
          const int i = g.idx[c0 + j];
          const FusedPrep<TW>& P = prep[i];
          const void* src = mode == ROW_MID ? in[i].U : in[i].u;
          void* dst = mode == ROW_LAST ? in[i].u : in[i].U;
          b.inst[j].in = reinterpret_cast<const cpx<float>*>(src);
          b.inst[j].out = reinterpret_cast<cpx<float>*>(dst);
          b.inst[j].ta = P.ta;
          P.at(b.inst[j].ta, first + k);
        }
        ok = fused_row_b<TW>(g.nx, g.rv, mode, g.tm, b, cnt, g.ny, prep[g.idx[c0]].tx, cs) &&
             hipGetLastError() == hipSuccess;
        ++kern;
      }
  };
  auto cols = [&](bool last) {
    for (auto& g : groups)
      for (size_t c0 = 0; c0 < g.idx.size() && ok; c0 += kMaxBatch) {
        const int cnt = (int)std::min<size_t>(kMaxBatch, g.idx.size() - c0);
        ColBatch b{};
        for (int j = 0; j < cnt; ++j) {
          const int i = g.idx[c0 + j];
          const ChirpTables* ct = prep[i].ct;
          b.inst[j] = ColInst{reinterpret_cast<const cpx<float>*>(in[i].U),
                              reinterpret_cast<cpx<float>*>(last ? in[i].u : in[i].U),
                              reinterpret_cast<cpx<float>*>(in[i].U), ct->Bx, ct->Ay, ct->Rmax};
        }
        const float spec_scale = (float)((double)g.nx * (double)g.ny);
        ok = fused_col_b<TW>(g.ny, g.cv, last && want_spec, b, cnt, g.nx, prep[g.idx[c0]].ty,
                             spec_scale, cs) &&
             hipGetLastError() == hipSuccess;
        ++kern;
      }
  };
  for (int k = 0; k < n && ok; ++k) {
    rows(k, k == 0 ? ROW_FIRST : ROW_MID);
    cols(k == n - 1);
  }
  if (ok) rows(n - 1, ROW_LAST);
  hipGraph_t g = nullptr;
  const hipError_t e2 = hipStreamEndCapture(cs, &g);
  ktime_suspend(false);
  if (!ok || e2 != hipSuccess || !g) {
    if (g) hipGraphDestroy(g);
    return fail(RSIM_E_HIP, "batch capture failed: %s", hipGetErrorString(ok ? e2 : hipGetLastError()));
  }
  *graph = g;
  *ngroups = (int)groups.size();
  *nkern = kern;
  return RSIM_OK;
}

int batch_create(rsim_batch** out, const rsim_batch_instance* inst, int count, int first, int n,
                 int flags, hipStream_t s) {
  // one arithmetic for the whole graph: compensated as soon as any instance
  // would need it
  const int asked = flags;
  for (int i = 0; i < count; ++i)
    flags |= default_math(inst[i].nx, inst[i].ny, n, asked) & RSIM_FLAG_MATH_COMP32;
  const bool want_spec = !(flags & RSIM_FLAG_NO_SPECTRUM);
  // the caller's pending work (uploads) completes before tables and tuning
  RSIM_HIP(hipStreamSynchronize(s));
  hipStream_t cs = nullptr;
  RSIM_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
  hipGraph_t graph = nullptr;
  int ng = 0, nk = 0;
  int rc;
  // every cached table the capture looks up is pinned for the batch's life
  std::vector<const void*> used;
  table_use_log(&used);
  if (flags & RSIM_FLAG_MATH_COMP32)
    rc = batch_capture<twc>(inst, count, first, n, want_spec, cs, &graph, &ng, &nk);
  else if (flags & RSIM_FLAG_MATH_FP64)
    rc = batch_capture<double>(inst, count, first, n, want_spec, cs, &graph, &ng, &nk);
  else
    rc = batch_capture<float>(inst, count, first, n, want_spec, cs, &graph, &ng, &nk);
  table_use_log(nullptr);
  hipStreamDestroy(cs);
  if (rc != RSIM_OK) return rc;
  rsim_batch* b = new rsim_batch();
  b->device = current_device();
  b->generation = cache_generation();
  for (const void* p : used)
    if (table_pin(p, +1)) b->pinned.push_back(p);
  b->graph = graph;
  b->groups = ng;
  b->kernels = nk;
  const hipError_t e = hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    hipGraphDestroy(graph);
    for (const void* p : b->pinned) table_pin(p, -1);
    delete b;
    return fail(RSIM_E_HIP, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
  }
  *out = b;
  return RSIM_OK;
}

int batch_run(rsim_batch* b, hipStream_t s) {
  if (current_device() != b->device)
    return fail(RSIM_E_INVALID_ARG, "batch belongs to device %d", b->device);
  if (b->generation != cache_generation())
    return fail(RSIM_E_INVALID_ARG,
                "batch was captured before rsim_release_caches freed its tables; create it again");
  RSIM_HIP(hipGraphLaunch(b->exec, s));
  return RSIM_OK;
}

int batch_info(const rsim_batch* b, int* groups, int* kernels) {
  if (groups) *groups = b->groups;
  if (kernels) *kernels = b->kernels;
  return RSIM_OK;
}

int batch_destroy(rsim_batch* b) {
  if (b->generation == cache_generation())
    for (const void* p : b->pinned) table_pin(p, -1);
  if (b->exec) hipGraphExecDestroy(b->exec);
  if (b->graph) hipGraphDestroy(b->graph);
  delete b;
  return RSIM_OK;
}

}  // namespace rsim
