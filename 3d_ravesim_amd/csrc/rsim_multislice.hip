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
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "rsim_fft_cfg.hpp"
#include "rsim_internal.hpp"
#include "rsim_tables.hpp"

namespace rsim {

int transmission(void* u, int nx, int ny, int dtype, double dx, double dy, double wl,
                 const rsim_volume* vol, int slice, double px, double py, double pz, double xpos,
                 double ypos, hipStream_t s);

// ---------------------------------------------------------------------------
// Fused kernels
// ---------------------------------------------------------------------------
struct TransArgsF {
  const uint8_t* ids;     // slice base (ny_g x nx_g)
  int nxg;
  AxisMapDev xm, ym;
  const float4* lut;      // per material (hi.re, hi.im, lo.re, lo.im)
  const float2* dbf;      // per material (delta, beta), bilinear mode
  float K;                // 2*pi*pz/wl
  int lut_mode;
};

__device__ __forceinline__ cpx<float> transmit(cpx<float> v, const TransArgsF& a, int x, int y0,
                                               float fy, bool iny) {
  if (!(iny && a.xm.inside[x])) return v;
  const int x0 = a.xm.i0[x];
  const uint8_t* r0 = a.ids + (long long)y0 * a.nxg + x0;
  if (a.lut_mode) {
    const float4 t = a.lut[r0[0]];
    // v * (hi + lo)
    const float rr = v.x * t.x - v.y * t.y + (v.x * t.z - v.y * t.w);
    const float ii = v.x * t.y + v.y * t.x + (v.x * t.w + v.y * t.z);
    return mk(rr, ii);
  }
  const uint8_t* r1 = r0 + a.nxg;
  const float fx = a.xm.fracf[x];
  const float2 d00 = a.dbf[r0[0]], d01 = a.dbf[r0[1]], d10 = a.dbf[r1[0]], d11 = a.dbf[r1[1]];
  const float gx = 1.f - fx, gy = 1.f - fy;
  const float de = (d00.x * gx + d01.x * fx) * gy + (d10.x * gx + d11.x * fx) * fy;
  const float be = (d00.y * gx + d01.y * fx) * gy + (d10.y * gx + d11.y * fx) * fy;
  const float e = expf(-a.K * be);
  float s, c;
  sincosf(a.K * de, &s, &c);
  const float tr = e * c, ti = -e * s;
  return mk(v.x * tr - v.y * ti, v.x * ti + v.y * tr);
}

enum { ROW_FIRST = 0, ROW_MID = 1, ROW_LAST = 2 };

template <class F, int C, int PS, bool SPLIT, int MODE>
__global__ void __launch_bounds__(F::TC * C) ms_row_kernel(const cpx<float>* in, cpx<float>* out,
                                                           int ny, uint32_t nbytes,
                                                           StageTw<float> tw, TransArgsF ta) {
  constexpr int N = F::N;
  __shared__ unsigned char lds[LdsSize<F, C, PS>::slots * (SPLIT ? 4 : 8)];
  const int tid = threadIdx.x;
  const int c = tid % C;
  const int t = tid / C;
  const int y = blockIdx.x * C + c;
  const bool live = y < ny;
  const int yy = live ? y : 0;
  const Gio<float> gi(in, nbytes), go(out, nbytes);
  const uint32_t vo = ((uint32_t)yy * N + t) * 8u;
  cpx<float> v[F::E];
#pragma unroll
  for (int e = 0; e < F::E; ++e) v[e] = gi.ld(vo, first_off<F>(e, 8u));
  if constexpr (MODE == ROW_FIRST) {
    const int y0 = ta.ym.i0[yy];
    const float fy = ta.ym.fracf[yy];
    const bool iny = ta.ym.inside[yy];
#pragma unroll
    for (int e = 0; e < F::E; ++e) v[e] = transmit(v[e], ta, first_pos<F>(t, e), y0, fy, iny);
    reg_fft<F, -1, C, PS, SPLIT>(v, tw, lds, t, c);
  } else if constexpr (MODE == ROW_MID) {
    reg_fft<F, 1, C, PS, SPLIT>(v, tw, lds, t, c);
    compiler_fence();
    const int y0 = ta.ym.i0[yy];
    const float fy = ta.ym.fracf[yy];
    const bool iny = ta.ym.inside[yy];
#pragma unroll
    for (int e = 0; e < F::E; ++e) v[e] = transmit(v[e], ta, last_pos<F>(t, e), y0, fy, iny);
    compiler_fence();
    reg_fft<F, -1, C, PS, SPLIT>(v, tw, lds, t, c);
  } else {
    reg_fft<F, 1, C, PS, SPLIT>(v, tw, lds, t, c);
  }
  if (live) {
#pragma unroll
    for (int e = 0; e < F::E; ++e) go.st(v[e], vo, last_off<F>(e, 8u));
  }
}

template <class F, int C, int PS, bool SPLIT, int SPEC>
__global__ void __launch_bounds__(F::TC * C) ms_col_kernel(const cpx<float>* in, cpx<float>* out,
                                                           cpx<float>* spec, int nx,
                                                           uint32_t nbytes, StageTw<float> tw,
                                                           const double2* __restrict__ Bx,
                                                           const double2* __restrict__ Ay,
                                                           const int* __restrict__ Rmax,
                                                           float spec_scale) {
  constexpr int N = F::N;
  __shared__ unsigned char lds[LdsSize<F, C, PS>::slots * (SPLIT ? 4 : 8)];
  const int tid = threadIdx.x;
  const int c = tid % C;
  const int t = tid / C;
  const int x = xcd_tile(blockIdx.x, gridDim.x) * C + c;
  const bool live = x < nx;
  const int xx = live ? x : 0;
  const Gio<float> gi(in, nbytes), go(out, nbytes), gs(spec, nbytes);
  const uint32_t row_b = (uint32_t)nx * 8u;
  const uint32_t vo = (uint32_t)t * row_b + (uint32_t)xx * 8u;
  cpx<float> v[F::E];
#pragma unroll
  for (int e = 0; e < F::E; ++e) v[e] = gi.ld(vo, first_off<F>(e, row_b));
  reg_fft<F, -1, C, PS, SPLIT>(v, tw, lds, t, c);
  compiler_fence();
  const double2 bx = Bx[xx];
  const int rmax = Rmax[xx];
  constexpr int half = (N - 1) / 2 + 1;
  // groups of 4 with a compiler fence in between bound the number of FP64
  // table values in flight (register pressure)
#pragma unroll
  for (int e = 0; e < F::E; ++e) {
    const int y = last_pos<F>(t, e);
    const int sy = y < half ? y : y - N;
    const int ay = sy < 0 ? -sy : sy;
    const double2 a = Ay[y];
    const double mr = bx.x * a.x - bx.y * a.y;
    const double mi = bx.x * a.y + bx.y * a.x;
    const double vr = (double)v[e].x, vi = (double)v[e].y;
    const cpx<float> w = mk((float)(vr * mr - vi * mi), (float)(vr * mi + vi * mr));
    v[e] = ay > rmax ? mk(0.f, 0.f) : w;
    if constexpr (SPEC) {
      if (live) gs.st(mk(v[e].x * spec_scale, v[e].y * spec_scale), vo, last_off<F>(e, row_b));
    }
    if ((e & 3) == 3) compiler_fence();
  }
  reg_fft<F, 1, C, PS, SPLIT>(v, tw, lds, t, c);
  if (live) {
#pragma unroll
    for (int e = 0; e < F::E; ++e) go.st(v[e], vo, last_off<F>(e, row_b));
  }
}

namespace {

template <int N>
struct FusedRow {
  using S = RowShape<float, N>;
  using F = typename S::F;
  static void launch(int mode, const void* in, void* out, int ny, const StageTw<float>& tw,
                     const TransArgsF& ta, hipStream_t s) {
    const uint32_t nb = (uint32_t)((long long)N * ny * 8);
    dim3 grid((ny + S::C - 1) / S::C), block(F::TC * S::C);
    const cpx<float>* i = (const cpx<float>*)in;
    cpx<float>* o = (cpx<float>*)out;
    if (mode == ROW_FIRST)
      hipLaunchKernelGGL((ms_row_kernel<F, S::C, S::PS, S::SPLIT, ROW_FIRST>), grid, block, 0, s, i, o, ny, nb, tw, ta);
    else if (mode == ROW_MID)
      hipLaunchKernelGGL((ms_row_kernel<F, S::C, S::PS, S::SPLIT, ROW_MID>), grid, block, 0, s, i, o, ny, nb, tw, ta);
    else
      hipLaunchKernelGGL((ms_row_kernel<F, S::C, S::PS, S::SPLIT, ROW_LAST>), grid, block, 0, s, i, o, ny, nb, tw, ta);
  }
};

template <int N>
struct FusedCol {
  using S = ColShape<float, N>;
  using F = typename S::F;
  static void launch(bool spec_out, const void* in, void* out, void* spec, int nx,
                     const StageTw<float>& tw, const double2* Bx, const double2* Ay,
                     const int* Rmax, float spec_scale, hipStream_t s) {
    dim3 grid((nx + S::C - 1) / S::C), block(F::TC * S::C);
    const uint32_t nb = (uint32_t)((long long)N * nx * 8);
    const cpx<float>* i = (const cpx<float>*)in;
    cpx<float>* o = (cpx<float>*)out;
    cpx<float>* sp = (cpx<float>*)spec;
    if (spec_out)
      hipLaunchKernelGGL((ms_col_kernel<F, S::C, S::PS, S::SPLIT, 1>), grid, block, 0, s, i, o, sp, nx, nb, tw, Bx, Ay, Rmax, spec_scale);
    else
      hipLaunchKernelGGL((ms_col_kernel<F, S::C, S::PS, S::SPLIT, 0>), grid, block, 0, s, i, o, sp, nx, nb, tw, Bx, Ay, Rmax, spec_scale);
  }
};

bool fused_row(int N, int mode, const void* in, void* out, int ny, const StageTw<float>& tw,
               const TransArgsF& ta, hipStream_t s) {
  switch (N) {
    case 256: FusedRow<256>::launch(mode, in, out, ny, tw, ta, s); return true;
    case 512: FusedRow<512>::launch(mode, in, out, ny, tw, ta, s); return true;
    case 1024: FusedRow<1024>::launch(mode, in, out, ny, tw, ta, s); return true;
    case 2048: FusedRow<2048>::launch(mode, in, out, ny, tw, ta, s); return true;
    case 4096: FusedRow<4096>::launch(mode, in, out, ny, tw, ta, s); return true;
    case 8192: FusedRow<8192>::launch(mode, in, out, ny, tw, ta, s); return true;
    default: return false;
  }
}

bool fused_col(int N, bool spec_out, const void* in, void* out, void* spec, int nx,
               const StageTw<float>& tw, const double2* Bx, const double2* Ay, const int* Rmax,
               float spec_scale, hipStream_t s) {
  switch (N) {
    case 256: FusedCol<256>::launch(spec_out, in, out, spec, nx, tw, Bx, Ay, Rmax, spec_scale, s); return true;
    case 512: FusedCol<512>::launch(spec_out, in, out, spec, nx, tw, Bx, Ay, Rmax, spec_scale, s); return true;
    case 1024: FusedCol<1024>::launch(spec_out, in, out, spec, nx, tw, Bx, Ay, Rmax, spec_scale, s); return true;
    case 2048: FusedCol<2048>::launch(spec_out, in, out, spec, nx, tw, Bx, Ay, Rmax, spec_scale, s); return true;
    case 4096: FusedCol<4096>::launch(spec_out, in, out, spec, nx, tw, Bx, Ay, Rmax, spec_scale, s); return true;
    case 8192: FusedCol<8192>::launch(spec_out, in, out, spec, nx, tw, Bx, Ay, Rmax, spec_scale, s); return true;
    default: return false;
  }
}

// FP64 chirp tables of the fused column pass.
struct ChirpTables {
  double2* Bx = nullptr;   // g * exp(i theta_x) / (nx*ny)
  double2* Ay = nullptr;   // exp(i theta_y)
  int* Rmax = nullptr;     // largest |ky index| kept for each column, -1: none
};

std::mutex g_mu;
std::map<std::tuple<int, int, int, uint64_t, uint64_t, uint64_t, uint64_t, uint64_t>, ChirpTables> g_chirp;

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

int chirp_tables(int nx, int ny, double dx, double dy, double wl, double dz, double freq,
                 const ChirpTables** out) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto key = std::make_tuple(current_device(), nx, ny, bits(dx), bits(dy), bits(wl), bits(dz), bits(freq));
  auto it = g_chirp.find(key);
  if (it != g_chirp.end()) {
    *out = &it->second;
    return RSIM_OK;
  }
  const std::vector<double> kx2 = host_k2(nx, dx), ky2 = host_k2(ny, dy);
  const double ctheta = ((-M_PI) * wl) * dz;
  const double gph = reduce_phase(((2.0 * M_PI) / wl) * dz);
  const double inv_n = 1.0 / ((double)nx * (double)ny);
  std::vector<double> bx(2 * nx), ay(2 * ny);
  std::vector<int> rmax(nx);
  for (int x = 0; x < nx; ++x) {
    const double th = reduce_phase(gph + reduce_phase(ctheta * kx2[x]));
    bx[2 * x] = std::cos(th) * inv_n;
    bx[2 * x + 1] = std::sin(th) * inv_n;
  }
  for (int y = 0; y < ny; ++y) {
    const double th = reduce_phase(ctheta * ky2[y]);
    ay[2 * y] = std::cos(th);
    ay[2 * y + 1] = std::sin(th);
  }
  const double fsq = 2.0 * (freq * freq);
  // ky2 by |index|: index m and -m have equal squares; -ny/2 exists for even ny
  for (int x = 0; x < nx; ++x) {
    int r = -1;
    for (int m = 0; m <= ny / 2; ++m) {
      const int yi = m;  // +m (or the Nyquist bin -ny/2 at index ny/2 for even ny)
      if (ky2[yi] + kx2[x] > fsq) break;
      r = m;
    }
    rmax[x] = r;
  }
  ChirpTables t;
  RSIM_HIP(hipMalloc((void**)&t.Bx, 16 * (size_t)nx));
  RSIM_HIP(hipMalloc((void**)&t.Ay, 16 * (size_t)ny));
  RSIM_HIP(hipMalloc((void**)&t.Rmax, 4 * (size_t)nx));
  RSIM_HIP(hipMemcpy(t.Bx, bx.data(), 16 * (size_t)nx, hipMemcpyHostToDevice));
  RSIM_HIP(hipMemcpy(t.Ay, ay.data(), 16 * (size_t)ny, hipMemcpyHostToDevice));
  RSIM_HIP(hipMemcpy(t.Rmax, rmax.data(), 4 * (size_t)nx, hipMemcpyHostToDevice));
  auto r = g_chirp.emplace(key, t);
  *out = &r.first->second;
  return RSIM_OK;
}

// Per-call material tables of the fused row pass (tiny; synchronous upload
// into a per-device buffer reused across calls).
struct MatBuf {
  void* dev = nullptr;
};
std::map<int, MatBuf> g_mat;

int material_tables(const rsim_volume* vol, double K, const float4** lut, const float2** dbf) {
  std::lock_guard<std::mutex> lk(g_mu);
  MatBuf& mb = g_mat[current_device()];
  if (!mb.dev) RSIM_HIP(hipMalloc(&mb.dev, 256 * (16 + 8)));
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

int release_multislice_tables() {
  std::lock_guard<std::mutex> lk(g_mu);
  for (auto& kv : g_chirp) {
    hipFree(kv.second.Bx);
    hipFree(kv.second.Ay);
    hipFree(kv.second.Rmax);
  }
  g_chirp.clear();
  for (auto& kv : g_mat) hipFree(kv.second.dev);
  g_mat.clear();
  return RSIM_OK;
}

bool fused_supported(int nx, int ny, int dtype) {
  SchedInfo a, b;
  return dtype == RSIM_C64 && sched_info(nx, &a) && sched_info(ny, &b);
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

int sample_apply(void* u, void* U, int nx, int ny, int dtype, double dx, double dy, double wl,
                 double freq, const rsim_volume* vol, double px, double py, double pz,
                 double xpos, double ypos, int first, int count, int engine, int flags,
                 hipStream_t s) {
  const int total = count;
  if (count == 0) return RSIM_OK;
  const bool want_spec = !(flags & RSIM_FLAG_NO_SPECTRUM);
  if (engine == RSIM_PLAIN || !fused_supported(nx, ny, dtype)) {
    for (int k = 0; k < count; ++k) {
      if (cancel_requested()) return fail(RSIM_E_CANCELLED, "multislice cancelled at slice %d", first + k);
      RSIM_CHECK(transmission(u, nx, ny, dtype, dx, dy, wl, vol, first + k, px, py, pz, xpos, ypos, s));
      RSIM_CHECK(propagate_plain(u, U, nx, ny, dtype, dx, dy, wl, pz, freq, false, s));
      if (g_progress_sync()) RSIM_HIP(hipStreamSynchronize(s));
      report_progress("sample_slice", k + 1, total);
    }
    return RSIM_OK;
  }
  // ---- fused complex64 engine ----
  const void* twx[4];
  const void* twy[4];
  if (!reg_stage_tables(dtype, nx, twx) || !reg_stage_tables(dtype, ny, twy))
    return fail(RSIM_E_HIP, "missing FFT schedule tables");
  StageTw<float> tx, ty;
  for (int i = 0; i < 4; ++i) {
    tx.t[i] = (const cpx<float>*)twx[i];
    ty.t[i] = (const cpx<float>*)twy[i];
  }
  const ChirpTables* ct = nullptr;
  RSIM_CHECK(chirp_tables(nx, ny, dx, dy, wl, pz, freq, &ct));
  const AxisMap* xm = nullptr;
  const AxisMap* ym = nullptr;
  RSIM_CHECK(axis_map(nx, dx, 1.0, xpos, vol->nx_g, px, &xm));
  RSIM_CHECK(axis_map(ny, dy, 1.0, ypos, vol->ny_g, py, &ym));
  const double K = ((2.0 * M_PI) * pz) / wl;
  TransArgsF ta;
  RSIM_CHECK(material_tables(vol, K, &ta.lut, &ta.dbf));
  ta.nxg = vol->nx_g;
  ta.xm = xm->dev;
  ta.ym = ym->dev;
  ta.K = (float)K;
  ta.lut_mode = (xm->all_integral && ym->all_integral) ? 1 : 0;
  const float spec_scale = (float)((double)nx * (double)ny);
  const size_t slice_bytes = (size_t)vol->ny_g * vol->nx_g;
  for (int k = 0; k < count; ++k) {
    const int sl = first + k;
    ta.ids = vol->ids + (size_t)sl * slice_bytes;
    const bool last = k == count - 1;
    ktime_begin(KT_ROW, s);
    if (!fused_row(nx, k == 0 ? ROW_FIRST : ROW_MID, k == 0 ? u : U, U, ny, tx, ta, s))
      return fail(RSIM_E_HIP, "no fused row kernel for nx=%d", nx);
    ktime_end(KT_ROW, s);
    RSIM_LAUNCHED();
    // last slice: W -> u (and the spectrum -> U); otherwise in place in U
    ktime_begin(KT_COL, s);
    if (!fused_col(ny, last && want_spec, U, last ? u : U, U, nx, ty, ct->Bx, ct->Ay, ct->Rmax,
                   spec_scale, s))
      return fail(RSIM_E_HIP, "no fused column kernel for ny=%d", ny);
    ktime_end(KT_COL, s);
    RSIM_LAUNCHED();
    const bool cancel = !last && cancel_requested();
    if (cancel || last) {
      // row IFFT of W; after a cancel this leaves u consistent at a slice
      // boundary (U then does not hold FFT2(u))
      ktime_begin(KT_ROW, s);
      fused_row(nx, ROW_LAST, last ? u : U, u, ny, tx, ta, s);
      ktime_end(KT_ROW, s);
      RSIM_LAUNCHED();
      if (cancel) return fail(RSIM_E_CANCELLED, "multislice cancelled after slice %d", sl);
    }
    if (g_progress_sync()) RSIM_HIP(hipStreamSynchronize(s));
    report_progress("sample_slice", k + 1, total);
  }
  return RSIM_OK;
}

}  // namespace rsim
