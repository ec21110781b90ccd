// PlasmaSample: slice transmissions from per-voxel plasma state
// (big-wave/plasma_sample.py:126-233, big-wave/plasma.py:380-434).
//
// The four grids (n_e, n_i, T_e, Z*) stay resident in HBM.  Per voxel the
// refractive decrement and absorption index are
//   delta = n_e P + [bound] n_i P (1-q) f1bar
//   beta  =         [bound] n_i P (1-q) f2eff_bar + [ff] alpha_ff lambda_cm / (4 pi)
//   P = r_e lambda_cm^2 / (2 pi),  q = clip(Z*/zbar, 0, 1),
//   alpha_ff = 3.7e8 n_e n_i q^2 z2bar / (sqrt(11604.5 T_e) nu^3)   (Kramers)
// evaluated in FP64 in the reference's operation order.  The plain engine
// computes one slice of (delta, beta) into a scratch slice and interpolates it
// bilinearly in FP64 (vacuum outside the grid, plasma_sample.py:172-218); the
// fused engine interpolates a cached FP32 (delta, beta) volume inside its row
// pass (transmission source TM_FIELD).
#include <atomic>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "rsim_device.hpp"
#include "rsim_internal.hpp"
#include "rsim_tables.hpp"

#pragma clang fp contract(off)

struct rsim_plasma {
  unsigned long long serial;
  int device;
  int nz, ny_g, nx_g;
  int elem;          // bytes per grid value: 4 (float32) or 8 (float64)
  void* grid[4];     // n_e [cm^-3], n_i [cm^-3], T_e [eV], Z*
};

namespace rsim {
namespace {

// nist_lookup/nist_lookup/physical_constants.py:20-27 as evaluated by the
// reference's pinned scipy (~=1.10, CODATA 2018): PLANCK_HC = 1e10 h c / e in
// eV*Angstrom, R_ELECTRON_CM = 100 r_e in cm.
const double kPlanckHc = ((1.0e10 * 6.62607015e-34) * 299792458.0) / 1.602176634e-19;
const double kRElectronCm = 100.0 * 2.8179403262e-15;

struct PlasmaK {
  double prefactor, f1bar, f2eff, zbar, z2bar, nu3, lamb_cm;
};

int plasma_consts(double energy, const double* optics, PlasmaK* k) {
  if (!(energy > 0.0) || !std::isfinite(energy))
    return fail(RSIM_E_INVALID_ARG, "energy must be positive, got %g", energy);
  if (!optics) return fail(RSIM_E_INVALID_ARG, "optics (f1bar, f2eff_bar, zbar, z2bar) is NULL");
  for (int i = 0; i < 4; ++i)
    if (!std::isfinite(optics[i])) return fail(RSIM_E_INVALID_ARG, "optics[%d] is not finite", i);
  k->lamb_cm = (1.0e-8 * kPlanckHc) / energy;                          // plasma.py:406
  k->prefactor = ((kRElectronCm * k->lamb_cm) * k->lamb_cm) / (2.0 * M_PI);  // plasma.py:407
  k->f1bar = optics[0];
  k->f2eff = optics[1];
  k->zbar = optics[2];
  k->z2bar = optics[3];
  const double nu = energy / 4.135667e-15;                             // plasma.py:419
  k->nu3 = std::pow(nu, 3.0);
  return RSIM_OK;
}

template <typename T>
__device__ __forceinline__ double2 plasma_db(const T* ne_g, const T* ni_g, const T* te_g,
                                             const T* zs_g, long long i, const PlasmaK& k) {
  const double ne = (double)ne_g[i], ni = (double)ni_g[i], te = (double)te_g[i];
  const double zs = (double)zs_g[i];
  double delta = ne * k.prefactor;
  double beta = 0.0;
  double q = k.zbar > 0.0 ? zs / k.zbar : 0.0;
  q = fmin(fmax(q, 0.0), 1.0);
  const double fb = 1.0 - q;
  if (fb > 0.0 && ni > 0.0) {
    const double bs = (ni * k.prefactor) * fb;
    delta = delta + bs * k.f1bar;
    beta = beta + bs * k.f2eff;
  }
  if (ne > 0.0 && ni > 0.0 && q > 0.0 && k.z2bar > 0.0 && te > 0.0) {
    const double alpha = (((3.7e8 * ne) * ni) * ((q * q) * k.z2bar)) / (sqrt(te * 11604.5) * k.nu3);
    beta = beta + (alpha * k.lamb_cm) / (4.0 * M_PI);
  }
  return make_double2(delta, beta);
}

template <typename T, typename O>
__global__ void __launch_bounds__(256) plasma_db_kernel(const T* __restrict__ ne,
                                                        const T* __restrict__ ni,
                                                        const T* __restrict__ te,
                                                        const T* __restrict__ zs, long long n,
                                                        PlasmaK k, O* __restrict__ out) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const double2 d = plasma_db(ne, ni, te, zs, i, k);
    if constexpr (sizeof(O) == 8)
      out[i] = make_float2((float)d.x, (float)d.y);
    else
      out[i] = d;
  }
}

// (delta, beta) of voxels [off, off + n) into out (double2 or float2).
template <typename O>
int plasma_db_launch(const rsim_plasma* p, long long off, long long n, const PlasmaK& k, O* out,
                     hipStream_t s) {
  const int blocks = (int)std::min<long long>((n + 255) / 256, 65536);
  if (p->elem == 8) {
    const double* const* g = reinterpret_cast<const double* const*>(p->grid);
    hipLaunchKernelGGL((plasma_db_kernel<double, O>), dim3(blocks), dim3(256), 0, s, g[0] + off,
                       g[1] + off, g[2] + off, g[3] + off, n, k, out);
  } else {
    const float* const* g = reinterpret_cast<const float* const*>(p->grid);
    hipLaunchKernelGGL((plasma_db_kernel<float, O>), dim3(blocks), dim3(256), 0, s, g[0] + off,
                       g[1] + off, g[2] + off, g[3] + off, n, k, out);
  }
  RSIM_LAUNCHED();
  return RSIM_OK;
}

// Plain-engine transmission from one (delta, beta) slice: the arithmetic of
// the material-id transmission kernel with the corners read from the field.
template <typename R>
__global__ void __launch_bounds__(256) field_transmission_kernel(cpx<R>* u, int nx, int ny,
                                                                 const double2* __restrict__ db,
                                                                 int nxg, const AxisMapDev xm,
                                                                 const AxisMapDev ym, double K) {
  const long long n = (long long)nx * ny;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < n;
       idx += (long long)gridDim.x * blockDim.x) {
    const int y = (int)(idx / nx);
    const int x = (int)(idx - (long long)y * nx);
    if (!(xm.inside[x] && ym.inside[y])) continue;   // vacuum: factor exactly 1
    const int x0 = xm.i0[x], y0 = ym.i0[y];
    const double fx = xm.frac[x], fy = ym.frac[y];
    const double2* r0 = db + (long long)y0 * nxg + x0;
    const double2* r1 = r0 + nxg;
    const double2 a = r0[0], b = r0[1], c = r1[0], d = r1[1];
    const double gx = 1.0 - fx, gy = 1.0 - fy;
    const double dy0r = a.x * gx + b.x * fx;
    const double dy0i = a.y * gx + b.y * fx;
    const double dy1r = c.x * gx + d.x * fx;
    const double dy1i = c.y * gx + d.y * fx;
    const double delta = dy0r * gy + dy1r * fy;
    const double beta = dy0i * gy + dy1i * fy;
    const double er = -(K * beta);
    const double ei = -(K * delta);
    const double e = exp(er);
    double sn, co;
    sincos(ei, &sn, &co);
    const double tr = e * co, ti = e * sn;
    const cpx<R> v = u[idx];
    const double vr = (double)v.x, vi = (double)v.y;
    u[idx] = mk((R)(vr * tr - vi * ti), (R)(vr * ti + vi * tr));
  }
}

int plasma_transmission(void* u, int nx, int ny, int dtype, double dx, double dy, double wl,
                        const rsim_plasma* p, int slice, const PlasmaK& k, double px, double py,
                        double pz, double xpos, double ypos, double scale, hipStream_t s) {
  const AxisMap* xm = nullptr;
  const AxisMap* ym = nullptr;
  RSIM_CHECK(axis_map(nx, dx, scale, xpos, p->nx_g, px, &xm));
  RSIM_CHECK(axis_map(ny, dy, scale, ypos, p->ny_g, py, &ym));
  const long long plane = (long long)p->ny_g * p->nx_g;
  double2* db = nullptr;
  RSIM_HIP(hipMallocAsync((void**)&db, sizeof(double2) * plane, s));
  int rc = plasma_db_launch<double2>(p, (long long)slice * plane, plane, k, db, s);
  if (rc == RSIM_OK) {
    const double K = ((2.0 * M_PI) * pz) / wl;   // plasma_sample.py:35-37
    const long long n = (long long)nx * ny;
    const int blocks = (int)std::min<long long>((n + 255) / 256, 65536);
    ktime_begin(KT_PW, s);
    if (dtype == RSIM_C64)
      hipLaunchKernelGGL(field_transmission_kernel<float>, dim3(blocks), dim3(256), 0, s,
                         (cpx<float>*)u, nx, ny, db, p->nx_g, xm->dev, ym->dev, K);
    else
      hipLaunchKernelGGL(field_transmission_kernel<double>, dim3(blocks), dim3(256), 0, s,
                         (cpx<double>*)u, nx, ny, db, p->nx_g, xm->dev, ym->dev, K);
    ktime_end(KT_PW, s);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) rc = fail(RSIM_E_HIP, "kernel launch failed: %s", hipGetErrorString(e));
  }
  const hipError_t ef = hipFreeAsync(db, s);
  if (rc == RSIM_OK && ef != hipSuccess)
    rc = fail(RSIM_E_HIP, "hipFreeAsync failed: %s", hipGetErrorString(ef));
  return rc;
}

// FP32 (delta, beta) volumes of the fused engine, one per (plasma, device,
// energy, optics), at most kFieldsPerPlasma per plasma (a multi-energy loop
// would otherwise keep one 8-byte-per-voxel volume per energy); freed with the
// plasma.  Evicting synchronises the device first, so no queued kernel still
// reads the volume.
constexpr int kFieldsPerPlasma = 2;
struct FieldEntry {
  void* dev = nullptr;
  unsigned long long stamp = 0;
};
std::mutex g_pmu;
unsigned long long g_field_clock = 0;
std::map<std::tuple<unsigned long long, int, uint64_t, uint64_t, uint64_t, uint64_t, uint64_t>,
         FieldEntry>
    g_fields;

uint64_t dbits(double v) {
  uint64_t b;
  std::memcpy(&b, &v, 8);
  return b;
}

int plasma_field(const rsim_plasma* p, double energy, const double* optics, const PlasmaK& k,
                 const float2** out, hipStream_t s) {
  std::lock_guard<std::mutex> lk(g_pmu);
  const auto key = std::make_tuple(p->serial, current_device(), dbits(energy), dbits(optics[0]),
                                   dbits(optics[1]), dbits(optics[2]), dbits(optics[3]));
  auto it = g_fields.find(key);
  if (it != g_fields.end()) {
    it->second.stamp = ++g_field_clock;
    *out = (const float2*)it->second.dev;
    return RSIM_OK;
  }
  // make room: drop this plasma's least recently used volumes
  for (;;) {
    int mine = 0;
    auto oldest = g_fields.end();
    for (auto e = g_fields.begin(); e != g_fields.end(); ++e) {
      if (std::get<0>(e->first) != p->serial) continue;
      ++mine;
      if (oldest == g_fields.end() || e->second.stamp < oldest->second.stamp) oldest = e;
    }
    if (mine < kFieldsPerPlasma) break;
    RSIM_HIP(hipDeviceSynchronize());
    (void)hipFree(oldest->second.dev);
    g_fields.erase(oldest);
  }
  const long long n = (long long)p->nz * p->ny_g * p->nx_g;
  void* f = nullptr;
  RSIM_HIP(hipMalloc(&f, sizeof(float2) * n));
  const int rc = plasma_db_launch<float2>(p, 0, n, k, (float2*)f, s);
  if (rc != RSIM_OK) {
    (void)hipFree(f);
    return rc;
  }
  g_fields[key] = FieldEntry{f, ++g_field_clock};
  *out = (const float2*)f;
  return RSIM_OK;
}

void plasma_release(unsigned long long serial) {
  std::lock_guard<std::mutex> lk(g_pmu);
  for (auto it = g_fields.begin(); it != g_fields.end();) {
    if (std::get<0>(it->first) == serial) {
      (void)hipFree(it->second.dev);
      it = g_fields.erase(it);
    } else {
      ++it;
    }
  }
}

int check_plasma(const rsim_plasma* p, int first, int n, bool interp = true) {
  if (!p) return fail(RSIM_E_INVALID_ARG, "plasma is NULL");
  if (interp && (p->ny_g < 2 || p->nx_g < 2))
    return fail(RSIM_E_INVALID_ARG, "a plasma sample needs at least 2x2 transverse voxels, got %dx%d",
                p->ny_g, p->nx_g);
  if (first < 0 || n < 0 || first + n > p->nz)
    return fail(RSIM_E_INVALID_ARG, "slices [%d, %d) outside the plasma's %d slices", first,
                first + n, p->nz);
  return RSIM_OK;
}

}  // namespace
}  // namespace rsim

using namespace rsim;

extern "C" {

int rsim_plasma_create(rsim_plasma** out, const void* ne, const void* ni, const void* te,
                       const void* zstar, int elem_size, int nz, int ny_g, int nx_g) {
  if (!out) return fail(RSIM_E_INVALID_ARG, "out is NULL");
  *out = nullptr;
  if (!ne || !ni || !te || !zstar) return fail(RSIM_E_INVALID_ARG, "plasma grid pointer is NULL");
  if (elem_size != 4 && elem_size != 8)
    return fail(RSIM_E_INVALID_ARG, "elem_size must be 4 (float32) or 8 (float64), got %d", elem_size);
  if (nz <= 0 || ny_g <= 0 || nx_g <= 0)
    return fail(RSIM_E_INVALID_ARG, "plasma grid dimensions must be positive, got (%d, %d, %d)",
                nz, ny_g, nx_g);
  static std::atomic<unsigned long long> serial{0};
  rsim_plasma* p = new rsim_plasma();
  p->serial = ++serial;
  p->device = current_device();
  p->nz = nz;
  p->ny_g = ny_g;
  p->nx_g = nx_g;
  p->elem = elem_size;
  const size_t bytes = (size_t)nz * ny_g * nx_g * elem_size;
  const void* src[4] = {ne, ni, te, zstar};
  for (int i = 0; i < 4; ++i) {
    hipError_t e = hipMalloc(&p->grid[i], bytes);
    if (e == hipSuccess) e = hipMemcpy(p->grid[i], src[i], bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      for (int j = 0; j <= i; ++j)
        if (p->grid[j]) (void)hipFree(p->grid[j]);
      delete p;
      return fail(RSIM_E_HIP, "plasma grid upload failed: %s", hipGetErrorString(e));
    }
  }
  *out = p;
  return RSIM_OK;
}

int rsim_plasma_destroy(rsim_plasma* p) {
  if (!p) return RSIM_OK;
  plasma_release(p->serial);
  for (int i = 0; i < 4; ++i)
    if (p->grid[i]) (void)hipFree(p->grid[i]);
  delete p;
  return RSIM_OK;
}

int rsim_plasma_deltabeta(const rsim_plasma* p, int first_slice, int n_slices, double energy,
                          const double* optics, void* out, void* stream) {
  RSIM_CHECK(check_plasma(p, first_slice, n_slices, false));
  if (!out) return fail(RSIM_E_INVALID_ARG, "out is NULL");
  PlasmaK k;
  RSIM_CHECK(plasma_consts(energy, optics, &k));
  const long long plane = (long long)p->ny_g * p->nx_g;
  if (n_slices == 0) return RSIM_OK;
  return plasma_db_launch<double2>(p, (long long)first_slice * plane, plane * n_slices, k,
                                   (double2*)out, resolve_stream(stream));
}

int rsim_plasma_transmission_2d(void* u, int nx, int ny, int dtype, double dx, double dy,
                                double wl, const rsim_plasma* p, int slice, double energy,
                                const double* optics, double pixel_size_x, double pixel_size_y,
                                double pixel_size_z, double x_pos, double y_pos,
                                double coord_scale, void* stream) {
  if (!u) return fail(RSIM_E_INVALID_ARG, "u is NULL");
  if (!valid_dtype(dtype)) return fail(RSIM_E_INVALID_ARG, "dtype must be RSIM_C64 or RSIM_C128");
  if (nx <= 0 || ny <= 0) return fail(RSIM_E_INVALID_ARG, "nx and ny must be positive");
  RSIM_CHECK(check_plasma(p, slice, 1));
  if (!(wl > 0) || !(pixel_size_x > 0) || !(pixel_size_y > 0) || !(pixel_size_z > 0))
    return fail(RSIM_E_INVALID_ARG, "wl and pixel sizes must be positive");
  if (!(coord_scale > 0) || !std::isfinite(coord_scale))
    return fail(RSIM_E_INVALID_ARG, "coord_scale must be positive and finite");
  PlasmaK k;
  RSIM_CHECK(plasma_consts(energy, optics, &k));
  return plasma_transmission(u, nx, ny, dtype, dx, dy, wl, p, slice, k, pixel_size_x, pixel_size_y,
                             pixel_size_z, x_pos, y_pos, coord_scale, resolve_stream(stream));
}

}  // extern "C"

namespace {
int plasma_apply(void* u, void* U, int nx, int ny, int dtype, double dx, double dy, double wl,
                 double cutoff_freq, const rsim_plasma* p, double energy, const double* optics,
                 double pixel_size_x, double pixel_size_y, double pixel_size_z, double x_pos,
                 double y_pos, int first_slice, int n_slices, const double* coord_scale,
                 const double* slice_dz, int engine, int flags, const rsim_history_frames* history,
                 void* stream) {
  if (!u || !U) return fail(RSIM_E_INVALID_ARG, "u and U must be non-NULL");
  if (!valid_dtype(dtype)) return fail(RSIM_E_INVALID_ARG, "dtype must be RSIM_C64 or RSIM_C128");
  if (nx <= 0 || ny <= 0) return fail(RSIM_E_INVALID_ARG, "nx and ny must be positive");
  if (!fft_size_supported(nx) || !fft_size_supported(ny))
    return fail(RSIM_E_INVALID_ARG, "FFT sizes must factor into 2, 3, 5: nx=%d, ny=%d", nx, ny);
  RSIM_CHECK(check_plasma(p, first_slice, n_slices));
  if (engine != RSIM_PLAIN && engine != RSIM_FUSED)
    return fail(RSIM_E_INVALID_ARG, "engine must be RSIM_PLAIN or RSIM_FUSED");
  if (!(wl > 0) || !(dx > 0) || !(dy > 0)) return fail(RSIM_E_INVALID_ARG, "dx, dy and wl must be positive");
  if (!(pixel_size_x > 0) || !(pixel_size_y > 0) || !(pixel_size_z > 0))
    return fail(RSIM_E_INVALID_ARG, "pixel sizes must be positive");
  for (int i = 0; i < n_slices; ++i) {
    if (coord_scale && (!(coord_scale[i] > 0) || !std::isfinite(coord_scale[i])))
      return fail(RSIM_E_INVALID_ARG, "coord_scale[%d] must be positive and finite", i);
    if (slice_dz && !std::isfinite(slice_dz[i]))
      return fail(RSIM_E_INVALID_ARG, "slice_dz[%d] must be finite", i);
  }
  PlasmaK k;
  RSIM_CHECK(plasma_consts(energy, optics, &k));
  HistSpec hs;
  if (history) {
    if (!history->frames && n_slices > 0)
      return fail(RSIM_E_INVALID_ARG, "history frames buffer is NULL");
    if (history->out_x <= 0 || history->out_y <= 0 || !(history->pixel_size_x > 0) ||
        !(history->pixel_size_y > 0) || !std::isfinite(history->z_start))
      return fail(RSIM_E_INVALID_ARG, "invalid history detector geometry");
    hs.ox = history->out_x;
    hs.oy = history->out_y;
    hs.dsx = history->pixel_size_x;
    hs.dsy = history->pixel_size_y;
    hs.z_start = history->z_start;
    hs.frames = history->frames;
  }
  hipStream_t s = resolve_stream(stream);
  VoxSrc src;
  src.nz = p->nz;
  src.ny_g = p->ny_g;
  src.nx_g = p->nx_g;
  if (engine == RSIM_FUSED && fused_supported(nx, ny, dtype))
    RSIM_CHECK(plasma_field(p, energy, optics, k, &src.field, s));
  src.plain = [&](void* f, int dt, int slice, double scale, hipStream_t st) {
    return plasma_transmission(f, nx, ny, dt, dx, dy, wl, p, slice, k, pixel_size_x, pixel_size_y,
                               pixel_size_z, x_pos, y_pos, scale, st);
  };
  return slices_apply(u, U, nx, ny, dtype, dx, dy, wl, cutoff_freq, src, pixel_size_x,
                      pixel_size_y, pixel_size_z, x_pos, y_pos, first_slice, n_slices, coord_scale,
                      slice_dz, engine, flags, s, history ? &hs : nullptr);
}
}  // namespace

extern "C" {

int rsim_plasma_apply_2d(void* u, void* U, int nx, int ny, int dtype, double dx, double dy,
                         double wl, double cutoff_freq, const rsim_plasma* p, double energy,
                         const double* optics, double pixel_size_x, double pixel_size_y,
                         double pixel_size_z, double x_pos, double y_pos, int first_slice,
                         int n_slices, const double* coord_scale, const double* slice_dz,
                         int engine, int flags, void* stream) {
  return plasma_apply(u, U, nx, ny, dtype, dx, dy, wl, cutoff_freq, p, energy, optics,
                      pixel_size_x, pixel_size_y, pixel_size_z, x_pos, y_pos, first_slice,
                      n_slices, coord_scale, slice_dz, engine, flags, nullptr, stream);
}

int rsim_plasma_apply_2d_history(void* u, void* U, int nx, int ny, int dtype, double dx,
                                 double dy, double wl, double cutoff_freq, const rsim_plasma* p,
                                 double energy, const double* optics, double pixel_size_x,
                                 double pixel_size_y, double pixel_size_z, double x_pos,
                                 double y_pos, int first_slice, int n_slices,
                                 const double* coord_scale, const double* slice_dz, int engine,
                                 int flags, const rsim_history_frames* history, void* stream) {
  if (!history) return fail(RSIM_E_INVALID_ARG, "history is NULL");
  return plasma_apply(u, U, nx, ny, dtype, dx, dy, wl, cutoff_freq, p, energy, optics,
                      pixel_size_x, pixel_size_y, pixel_size_z, x_pos, y_pos, first_slice,
                      n_slices, coord_scale, slice_dz, engine, flags, history, stream);
}

}  // extern "C"
