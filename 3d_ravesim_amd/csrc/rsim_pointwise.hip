// Pointwise and reduction kernels of the plain engine, evaluated in FP64 in
// the reference's operation order (contraction disabled) so the complex128
// path tracks big-wave's NumpyVector arithmetic to round-off.
#include <algorithm>
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

namespace rsim {

// ---------------------------------------------------------------------------
// Chirp + circular cutoff (big-wave/propagation.py:396-442, 298-326)
// ---------------------------------------------------------------------------
template <typename R>
__global__ void __launch_bounds__(256) chirp_cutoff_kernel(cpx<R>* U, int nx, int ny,
                                                           const double* __restrict__ kx2,
                                                           const double* __restrict__ ky2,
                                                           double fsq, int chirp, double ctheta,
                                                           double gr, double gi) {
  const long long n = (long long)nx * ny;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < n;
       idx += (long long)gridDim.x * blockDim.x) {
    const int y = (int)(idx / nx);
    const int x = (int)(idx - (long long)y * nx);
    const double k2 = ky2[y] + kx2[x];
    cpx<R> v = U[idx];
    if (k2 > fsq) {
      v = mk(R(0), R(0));
    } else if (chirp) {
      double s, c;
      sincos(ctheta * k2, &s, &c);
      // h = global_phase * exp(i*theta)
      const double hr = gr * c - gi * s;
      const double hi = gr * s + gi * c;
      const double vr = (double)v.x, vi = (double)v.y;
      v = mk((R)(vr * hr - vi * hi), (R)(vr * hi + vi * hr));
    }
    U[idx] = v;
  }
}

int chirp_cutoff(void* U, int nx, int ny, int dtype, double dx, double dy, double wl, double dz,
                 double freq, bool chirp, hipStream_t s) {
  const AxisK2* kx = nullptr;
  const AxisK2* ky = nullptr;
  RSIM_CHECK(axis_k2(nx, dx, &kx));
  RSIM_CHECK(axis_k2(ny, dy, &ky));
  const double fsq = 2.0 * (freq * freq);
  // -1j*pi*wl*dz (propagation.py:430) and exp(1j*2*pi/wl*dz) (:429)
  const double ctheta = ((-M_PI) * wl) * dz;
  const double gph = ((2.0 * M_PI) / wl) * dz;
  const double gr = std::cos(gph), gi = std::sin(gph);
  const long long n = (long long)nx * ny;
  int blocks = (int)std::min<long long>((n + 255) / 256, 65536);
  ktime_begin(KT_PW, s);
  if (dtype == RSIM_C64)
    hipLaunchKernelGGL(chirp_cutoff_kernel<float>, dim3(blocks), dim3(256), 0, s, (cpx<float>*)U,
                       nx, ny, kx->dev, ky->dev, fsq, chirp ? 1 : 0, ctheta, gr, gi);
  else
    hipLaunchKernelGGL(chirp_cutoff_kernel<double>, dim3(blocks), dim3(256), 0, s,
                       (cpx<double>*)U, nx, ny, kx->dev, ky->dev, fsq, chirp ? 1 : 0, ctheta, gr,
                       gi);
  ktime_end(KT_PW, s);
  RSIM_LAUNCHED();
  return RSIM_OK;
}

// ---------------------------------------------------------------------------
// Point source (big-wave/propagation.py:458-484)
// ---------------------------------------------------------------------------
template <typename R>
__global__ void __launch_bounds__(256) point_source_kernel(cpx<R>* u, int nx, int ny, double dx,
                                                           double dy, double xs, double ys,
                                                           double z2, double k) {
  const long long n = (long long)nx * ny;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < n;
       idx += (long long)gridDim.x * blockDim.x) {
    const int y = (int)(idx / nx);
    const int x = (int)(idx - (long long)y * nx);
    // x_i = (i - nx/2) dx - xs, squared, in the reference's operation order
    // (no contraction in this file)
    const double vx = ((double)x - nx / 2.0) * dx - xs;
    const double vy = ((double)y - ny / 2.0) * dy - ys;
    const double r = sqrt(vy * vy + vx * vx + z2);
    double s, c;
    sincos(r * k, &s, &c);
    u[idx] = mk((R)(c / r), (R)(s / r));
  }
}

// ---------------------------------------------------------------------------
// Tilted plane wave of the Fresnel-scaled point source (big-wave/source.py:
// 62-90): u = exp(i phase) / dz, phase = -k (x xs + y ys) / dz, FP64.
// ---------------------------------------------------------------------------
template <typename R>
__global__ void __launch_bounds__(256) tilted_plane_kernel(cpx<R>* u, int nx, int ny, double dx,
                                                           double dy, double k, double xs,
                                                           double ys, double dz) {
  const long long n = (long long)nx * ny;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < n;
       idx += (long long)gridDim.x * blockDim.x) {
    const int iy = (int)(idx / nx);
    const int ix = (int)(idx - (long long)iy * nx);
    const double x = ((double)ix - nx / 2.0) * dx;
    const double y = ((double)iy - ny / 2.0) * dy;
    const double phase = ((-k) * (x * xs + y * ys)) / dz;
    double sn, cs;
    sincos(phase, &sn, &cs);
    u[idx] = mk((R)(cs / dz), (R)(sn / dz));
  }
}

int tilted_plane(void* u, int nx, int ny, int dtype, double dx, double dy, double wl, double dz,
                 double xs, double ys, hipStream_t s) {
  const double k = (2.0 * M_PI) / wl;
  const long long n = (long long)nx * ny;
  const int blocks = (int)std::min<long long>((n + 255) / 256, 65536);
  if (dtype == RSIM_C64)
    hipLaunchKernelGGL(tilted_plane_kernel<float>, dim3(blocks), dim3(256), 0, s, (cpx<float>*)u,
                       nx, ny, dx, dy, k, xs, ys, dz);
  else
    hipLaunchKernelGGL(tilted_plane_kernel<double>, dim3(blocks), dim3(256), 0, s,
                       (cpx<double>*)u, nx, ny, dx, dy, k, xs, ys, dz);
  RSIM_LAUNCHED();
  return RSIM_OK;
}

int point_source(void* u, int nx, int ny, int dtype, double dx, double dy, double z, double xs,
                 double ys, double wl, hipStream_t s) {
  const double k = (2.0 * M_PI) / wl;
  const long long n = (long long)nx * ny;
  int blocks = (int)std::min<long long>((n + 255) / 256, 65536);
  if (dtype == RSIM_C64)
    hipLaunchKernelGGL(point_source_kernel<float>, dim3(blocks), dim3(256), 0, s,
                       (cpx<float>*)u, nx, ny, dx, dy, xs, ys, z * z, k);
  else
    hipLaunchKernelGGL(point_source_kernel<double>, dim3(blocks), dim3(256), 0, s,
                       (cpx<double>*)u, nx, ny, dx, dy, xs, ys, z * z, k);
  RSIM_LAUNCHED();
  return RSIM_OK;
}

// ---------------------------------------------------------------------------
// Sample transmission, plain engine (big-wave/optical_element.py:538-584)
// ---------------------------------------------------------------------------
template <typename R>
__global__ void __launch_bounds__(256) transmission_kernel(cpx<R>* u, int nx, int ny,
                                                           const uint8_t* __restrict__ ids,
                                                           int nxg, const AxisMapDev xm,
                                                           const AxisMapDev ym,
                                                           const double* __restrict__ db,
                                                           double K) {
  const long long n = (long long)nx * ny;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < n;
       idx += (long long)gridDim.x * blockDim.x) {
    const int y = (int)(idx / nx);
    const int x = (int)(idx - (long long)y * nx);
    if (!(xm.inside[x] && ym.inside[y])) continue;   // vacuum: factor exactly 1
    const int x0 = xm.i0[x], y0 = ym.i0[y];
    const double fx = xm.frac[x], fy = ym.frac[y];
    const uint8_t* r0 = ids + (long long)y0 * nxg + x0;
    const uint8_t* r1 = r0 + nxg;
    const int a = r0[0], b = r0[1], c = r1[0], d = r1[1];
    const double gx = 1.0 - fx, gy = 1.0 - fy;
    const double dy0r = db[2 * a] * gx + db[2 * b] * fx;
    const double dy0i = db[2 * a + 1] * gx + db[2 * b + 1] * fx;
    const double dy1r = db[2 * c] * gx + db[2 * d] * fx;
    const double dy1i = db[2 * c + 1] * gx + db[2 * d + 1] * fx;
    const double delta = dy0r * gy + dy1r * fy;
    const double beta = dy0i * gy + dy1i * fy;
    const double er = -(K * beta);
    const double ei = -(K * delta);
    const double e = exp(er);
    double s, co;
    sincos(ei, &s, &co);
    const double tr = e * co, ti = e * s;
    const cpx<R> v = u[idx];
    const double vr = (double)v.x, vi = (double)v.y;
    u[idx] = mk((R)(vr * tr - vi * ti), (R)(vr * ti + vi * tr));
  }
}

int transmission(void* u, int nx, int ny, int dtype, double dx, double dy, double wl,
                 const rsim_volume* vol, int slice, double px, double py, double pz, double xpos,
                 double ypos, double coord_scale, hipStream_t s) {
  const AxisMap* xm = nullptr;
  const AxisMap* ym = nullptr;
  RSIM_CHECK(axis_map(nx, dx, coord_scale, xpos, vol->nx_g, px, &xm));
  RSIM_CHECK(axis_map(ny, dy, coord_scale, ypos, vol->ny_g, py, &ym));
  const double* db = nullptr;
  RSIM_CHECK(volume_db_dev(vol, &db));
  // -2j*pi*thickness/wl (optical_element.py:42)
  const double K = ((2.0 * M_PI) * pz) / wl;
  const uint8_t* ids = vol->ids + (long long)slice * vol->ny_g * vol->nx_g;
  const long long n = (long long)nx * ny;
  int blocks = (int)std::min<long long>((n + 255) / 256, 65536);
  ktime_begin(KT_PW, s);
  if (dtype == RSIM_C64)
    hipLaunchKernelGGL(transmission_kernel<float>, dim3(blocks), dim3(256), 0, s,
                       (cpx<float>*)u, nx, ny, ids, vol->nx_g, xm->dev, ym->dev, db, K);
  else
    hipLaunchKernelGGL(transmission_kernel<double>, dim3(blocks), dim3(256), 0, s,
                       (cpx<double>*)u, nx, ny, ids, vol->nx_g, xm->dev, ym->dev, db, K);
  ktime_end(KT_PW, s);
  RSIM_LAUNCHED();
  return RSIM_OK;
}

// ---------------------------------------------------------------------------
// area_v1 detector (big-wave/propagation.py:764-816, 949-1110)
// ---------------------------------------------------------------------------
struct DetPlanDev {
  const int* xi0;      // per output column
  const int* xi1;
  const double* xf0;
  const double* xf1;
  const double* xdet;  // detector x centres
  const int* jlo;      // per output row: first / one-past-last field row
  const int* jhi;
  const double* ylo;   // detector row edges
  const double* yhi;
  const double* ydet;
};

template <typename R>
__global__ void __launch_bounds__(256) area_v1_kernel(const cpx<R>* __restrict__ u, int nx,
                                                      int ny, int ox, int oy, DetPlanDev pl,
                                                      double dx, double dy, double gyl,
                                                      double z, double weight, int accumulate,
                                                      double* out) {
  const int p = blockIdx.x * 16 + threadIdx.x % 16;
  const int q = blockIdx.y * 16 + threadIdx.x / 16;
  if (p >= ox || q >= oy) return;
  const int i0 = pl.xi0[p], i1 = pl.xi1[p];
  const double f0 = pl.xf0[p], f1 = pl.xf1[p];
  const double ylo = pl.ylo[q], yhi = pl.yhi[q];
  const double xd = pl.xdet[p], yd = pl.ydet[q];
  const double cosine = z / sqrt(xd * xd + yd * yd + z * z);
  double acc = 0.0;
  for (int j = pl.jlo[q]; j < pl.jhi[q]; ++j) {
    const double cell_lo = gyl + j * dy;
    const double cell_hi = cell_lo + dy;
    const double ov = fmin(cell_hi, yhi) - fmax(cell_lo, ylo);
    if (!(ov > 0.0)) continue;
    const cpx<R>* row = u + (long long)j * nx;
    double s = 0.0;
    for (int i = i0; i < i1; ++i) {
      const cpx<R> v = row[i];
      s += (double)(v.x * v.x + v.y * v.y);
    }
    double xi = s * dx;
    if (i1 < nx) {
      const cpx<R> v = row[i1];
      xi += f1 * dx * (double)(v.x * v.x + v.y * v.y);
    }
    if (i0 < nx) {
      const cpx<R> v = row[i0];
      xi -= f0 * dx * (double)(v.x * v.x + v.y * v.y);
    }
    acc += xi * ov * cosine;
  }
  const long long o = (long long)q * ox + p;
  if (accumulate) out[o] += weight * acc;
  else out[o] = acc;
}

// Separable form of area_v1_kernel, same arithmetic in the same order.
// Stage 1: per field row j and detector column p, the x-integral
//   xs[j - j0][p] = dx * (sum_{i0 <= i < i1} |u|^2 + f1 |u[i1]|^2 - f0 |u[i0]|^2).
// xi0/xi1 are non-decreasing in p, so the 256 pixels of a block read one
// contiguous span of the row: it is loaded once, coalesced, as |u|^2 into LDS
// and the per-pixel sums run from there (directly from the row when the span
// exceeds kDetSpan cells).
constexpr int kDetSpan = 4096;
template <typename R>
__global__ void __launch_bounds__(256) area_v1_rows_kernel(const cpx<R>* __restrict__ u, int nx,
                                                           int ox, int j0, int nj, DetPlanDev pl,
                                                           double dx, double* __restrict__ xs) {
  __shared__ R e2[kDetSpan];
  const int pb = blockIdx.x * 256;
  const int jj = blockIdx.y;
  if (pb >= ox || jj >= nj) return;
  const int p = pb + threadIdx.x;
  const int pe = min(pb + 256, ox) - 1;
  const cpx<R>* row = u + (long long)(j0 + jj) * nx;
  const int lo = min(pl.xi0[pb], nx);
  const int hi = min(pl.xi1[pe] + 1, nx);
  if (hi - lo <= kDetSpan) {
    // four loads in flight per thread (clamped indexes stay in the span)
    for (int k0 = lo + (int)threadIdx.x; k0 < hi; k0 += 1024) {
      cpx<R> q[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) q[k] = row[min(k0 + k * 256, hi - 1)];
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (k0 + k * 256 < hi) e2[k0 + k * 256 - lo] = q[k].x * q[k].x + q[k].y * q[k].y;
    }
    __syncthreads();
    if (p >= ox) return;
    const int i0 = pl.xi0[p], i1 = pl.xi1[p];
    double s = 0.0;
    for (int i = i0; i < i1; ++i) s += (double)e2[i - lo];
    double xi = s * dx;
    if (i1 < nx) xi += pl.xf1[p] * dx * (double)e2[i1 - lo];
    if (i0 < nx) xi -= pl.xf0[p] * dx * (double)e2[i0 - lo];
    xs[(long long)jj * ox + p] = xi;
    return;
  }
  if (p >= ox) return;
  const int i0 = pl.xi0[p], i1 = pl.xi1[p];
  double s = 0.0;
  for (int i = i0; i < i1; ++i) {
    const cpx<R> v = row[i];
    s += (double)(v.x * v.x + v.y * v.y);
  }
  double xi = s * dx;
  if (i1 < nx) {
    const cpx<R> v = row[i1];
    xi += pl.xf1[p] * dx * (double)(v.x * v.x + v.y * v.y);
  }
  if (i0 < nx) {
    const cpx<R> v = row[i0];
    xi -= pl.xf0[p] * dx * (double)(v.x * v.x + v.y * v.y);
  }
  xs[(long long)jj * ox + p] = xi;
}

// Stage 2: per detector pixel, the y-overlap sum of the stage-1 rows times
// the obliquity factor.
__global__ void __launch_bounds__(256) area_v1_cols_kernel(const double* __restrict__ xs, int ox,
                                                           int oy, int j0, DetPlanDev pl,
                                                           double dy, double gyl, double z,
                                                           double weight, int accumulate,
                                                           double* out) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  const int q = blockIdx.y;
  if (p >= ox || q >= oy) return;
  const double ylo = pl.ylo[q], yhi = pl.yhi[q];
  const double xd = pl.xdet[p], yd = pl.ydet[q];
  const double cosine = z / sqrt(xd * xd + yd * yd + z * z);
  double acc = 0.0;
  // unconditional loads (rows [jlo, jhi) all lie in the scratch) so that
  // the unrolled iterations overlap them
#pragma unroll 4
  for (int j = pl.jlo[q]; j < pl.jhi[q]; ++j) {
    const double cell_lo = gyl + j * dy;
    const double cell_hi = cell_lo + dy;
    const double ov = fmin(cell_hi, yhi) - fmax(cell_lo, ylo);
    const double x = xs[(long long)(j - j0) * ox + p];
    if (ov > 0.0) acc += x * ov * cosine;
  }
  const long long o = (long long)q * ox + p;
  if (accumulate) out[o] += weight * acc;
  else out[o] = acc;
}

namespace {

// Host plan of the per-axis overlaps (propagation.py:949-977).
struct DetPlanHost {
  std::vector<int> xi0, xi1, jlo, jhi;
  std::vector<double> xf0, xf1, xdet, ylo, yhi, ydet;
  double gyl;
};

void det_plan(int nx, int ny, double dx, double dy, int ox, int oy, double dsx, double dsy,
              DetPlanHost* P) {
  P->xdet.resize(ox);
  P->ydet.resize(oy);
  for (int p = 0; p < ox; ++p) P->xdet[p] = ((double)p - (double)(ox / 2)) * dsx;
  for (int q = 0; q < oy; ++q) P->ydet[q] = ((double)q - (double)(oy / 2)) * dsy;
  P->xi0.resize(ox); P->xi1.resize(ox); P->xf0.resize(ox); P->xf1.resize(ox);
  P->jlo.resize(oy); P->jhi.resize(oy); P->ylo.resize(oy); P->yhi.resize(oy);
  const double gl = (-(double)(nx / 2) - 0.5) * dx;
  const double gr = gl + nx * dx;
  for (int p = 0; p < ox; ++p) {
    const double lo = P->xdet[p] - dsx * 0.5, hi = P->xdet[p] + dsx * 0.5;
    const double clo = std::min(std::max(lo, gl), gr);
    double chi = std::min(std::max(hi, gl), gr);
    chi = std::max(chi, clo);
    const double slo = (clo - gl) / dx, shi = (chi - gl) / dx;
    long long a = (long long)std::floor(slo), b = (long long)std::floor(shi);
    double fa = slo - (double)a, fb = shi - (double)b;
    if (a >= nx) { a = nx; fa = 0.0; }
    if (b >= nx) { b = nx; fb = 0.0; }
    P->xi0[p] = (int)a; P->xi1[p] = (int)b; P->xf0[p] = fa; P->xf1[p] = fb;
  }
  P->gyl = (-(double)(ny / 2) - 0.5) * dy;
  for (int q = 0; q < oy; ++q) {
    P->ylo[q] = P->ydet[q] - dsy * 0.5;
    P->yhi[q] = P->ydet[q] + dsy * 0.5;
    // rows whose cell [gyl + j dy, gyl + j dy + dy] can overlap (checked exactly in the kernel)
    long long a = (long long)std::floor((P->ylo[q] - P->gyl) / dy) - 1;
    long long b = (long long)std::floor((P->yhi[q] - P->gyl) / dy) + 2;
    a = std::max<long long>(0, std::min<long long>(a, ny));
    b = std::max<long long>(0, std::min<long long>(b, ny));
    P->jlo[q] = (int)a;
    P->jhi[q] = (int)std::max(a, b);
  }
}

// One device blob: int arrays, then (8-byte aligned) double arrays.
struct Blob {
  std::vector<unsigned char> bytes;
  std::vector<size_t> ioff, doff;
};
Blob make_blob(const std::vector<const std::vector<int>*>& iv,
               const std::vector<const std::vector<double>*>& dv) {
  Blob b;
  size_t o = 0;
  for (auto* v : iv) { b.ioff.push_back(o); o += v->size() * 4; }
  o = (o + 7) / 8 * 8;
  for (auto* v : dv) { b.doff.push_back(o); o += v->size() * 8; }
  b.bytes.resize(std::max<size_t>(o, 8));
  for (size_t k = 0; k < iv.size(); ++k)
    if (!iv[k]->empty()) std::memcpy(b.bytes.data() + b.ioff[k], iv[k]->data(), iv[k]->size() * 4);
  for (size_t k = 0; k < dv.size(); ++k)
    if (!dv[k]->empty()) std::memcpy(b.bytes.data() + b.doff[k], dv[k]->data(), dv[k]->size() * 8);
  return b;
}

int check_detector(int ox, int oy, double z) {
  if (ox <= 0 || oy <= 0) return fail(RSIM_E_INVALID_ARG, "detector output must be non-empty");
  if (!std::isfinite(z) || z == 0.0)
    return fail(RSIM_E_INVALID_ARG, "detector current_z must be finite and non-zero");
  return RSIM_OK;
}

}  // namespace

namespace {

// Device plans, one per (device, field grid, detector grid): the detector
// distance only enters the kernel's obliquity factor, so a history of frames
// at many z shares one plan.  Built synchronously once, kept until released.
struct DetCached {
  unsigned char* d = nullptr;
  DetPlanDev pl{};
  double gyl = 0.0;
  int j0 = 0, nj = 0;   // field rows any detector row reads: [j0, j0 + nj)
  bool mono = false;    // xi0, xi1 non-decreasing in p (the separable form needs it)
};
std::mutex g_det_mu;
std::map<std::tuple<int, int, int, uint64_t, uint64_t, int, int, uint64_t, uint64_t>, DetCached>
    g_det;

uint64_t fbits(double v) {
  uint64_t b;
  std::memcpy(&b, &v, 8);
  return b;
}

// Point c's plan at the blob b of P uploaded to d, and derive the row span
// and the monotonicity the two-stage form needs.
void det_plan_bind(const DetPlanHost& P, const Blob& b, unsigned char* d, int ox, int oy, int ny,
                   DetCached* c) {
  c->d = d;
  c->pl.xi0 = (const int*)(d + b.ioff[0]); c->pl.xi1 = (const int*)(d + b.ioff[1]);
  c->pl.jlo = (const int*)(d + b.ioff[2]); c->pl.jhi = (const int*)(d + b.ioff[3]);
  c->pl.xf0 = (const double*)(d + b.doff[0]); c->pl.xf1 = (const double*)(d + b.doff[1]);
  c->pl.xdet = (const double*)(d + b.doff[2]); c->pl.ylo = (const double*)(d + b.doff[3]);
  c->pl.yhi = (const double*)(d + b.doff[4]); c->pl.ydet = (const double*)(d + b.doff[5]);
  c->gyl = P.gyl;
  int jmin = ny, jmax = 0;
  for (int q = 0; q < oy; ++q)
    if (P.jhi[q] > P.jlo[q]) {
      jmin = std::min(jmin, P.jlo[q]);
      jmax = std::max(jmax, P.jhi[q]);
    }
  c->j0 = jmax > jmin ? jmin : 0;
  c->nj = jmax > jmin ? jmax - jmin : 0;
  c->mono = true;
  for (int p = 1; p < ox; ++p)
    if (P.xi0[p] < P.xi0[p - 1] || P.xi1[p] < P.xi1[p - 1]) c->mono = false;
}

bool det_two_stage(const DetCached& c, int oy) {
  return c.mono && c.nj > 0 && c.nj <= 65535 && oy <= 65535;
}

// The detector's launches: the separable two-stage form through xs (nj x ox
// doubles) when given, else the one-pass kernel.
int det_launch(const void* u, int nx, int ny, int dtype, double dx, double dy, int ox, int oy,
               double z, double weight, bool accumulate, double* out_dev, const DetCached& c,
               double* xs, hipStream_t s) {
  if (xs) {
    const dim3 g1((ox + 255) / 256, c.nj), g2((ox + 255) / 256, oy);
    if (dtype == RSIM_C64)
      hipLaunchKernelGGL(area_v1_rows_kernel<float>, g1, dim3(256), 0, s, (const cpx<float>*)u,
                         nx, ox, c.j0, c.nj, c.pl, dx, xs);
    else
      hipLaunchKernelGGL(area_v1_rows_kernel<double>, g1, dim3(256), 0, s,
                         (const cpx<double>*)u, nx, ox, c.j0, c.nj, c.pl, dx, xs);
    hipLaunchKernelGGL(area_v1_cols_kernel, g2, dim3(256), 0, s, xs, ox, oy, c.j0, c.pl, dy,
                       c.gyl, z, weight, accumulate ? 1 : 0, out_dev);
  } else {
    dim3 grid((ox + 15) / 16, (oy + 15) / 16);
    if (dtype == RSIM_C64)
      hipLaunchKernelGGL(area_v1_kernel<float>, grid, dim3(256), 0, s, (const cpx<float>*)u, nx,
                         ny, ox, oy, c.pl, dx, dy, c.gyl, z, weight, accumulate ? 1 : 0, out_dev);
    else
      hipLaunchKernelGGL(area_v1_kernel<double>, grid, dim3(256), 0, s, (const cpx<double>*)u,
                         nx, ny, ox, oy, c.pl, dx, dy, c.gyl, z, weight, accumulate ? 1 : 0,
                         out_dev);
  }
  RSIM_LAUNCHED();
  return RSIM_OK;
}

int det_plan_dev(int nx, int ny, double dx, double dy, int ox, int oy, double dsx, double dsy,
                 const DetCached** out) {
  std::lock_guard<std::mutex> lk(g_det_mu);
  const auto key = std::make_tuple(current_device(), nx, ny, fbits(dx), fbits(dy), ox, oy,
                                   fbits(dsx), fbits(dsy));
  auto it = g_det.find(key);
  if (it != g_det.end()) {
    *out = &it->second;
    return RSIM_OK;
  }
  RSIM_CHECK(capture_miss("detector plans"));
  DetPlanHost P;
  det_plan(nx, ny, dx, dy, ox, oy, dsx, dsy, &P);
  Blob b = make_blob({&P.xi0, &P.xi1, &P.jlo, &P.jhi},
                     {&P.xf0, &P.xf1, &P.xdet, &P.ylo, &P.yhi, &P.ydet});
  DetCached c;
  RSIM_HIP(hipMalloc((void**)&c.d, b.bytes.size()));
  const hipError_t e = hipMemcpy(c.d, b.bytes.data(), b.bytes.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(c.d);
    return fail(RSIM_E_HIP, "detector plan upload failed: %s", hipGetErrorString(e));
  }
  det_plan_bind(P, b, c.d, ox, oy, ny, &c);
  *out = &(g_det[key] = c);
  return RSIM_OK;
}

}  // namespace

int release_detector_plans() {
  std::lock_guard<std::mutex> lk(g_det_mu);
  for (auto& kv : g_det) (void)hipFree(kv.second.d);
  g_det.clear();
  return RSIM_OK;
}

int detector_area_v1(const void* u, int nx, int ny, int dtype, double dx, double dy, int ox,
                     int oy, double dsx, double dsy, double z, double weight, bool accumulate,
                     double* out_dev, hipStream_t s) {
  RSIM_CHECK(check_detector(ox, oy, z));
  const DetCached* c = nullptr;
  RSIM_CHECK(det_plan_dev(nx, ny, dx, dy, ox, oy, dsx, dsy, &c));
  // separable two-stage form through a stream-ordered scratch of the
  // x-integrals (nj x ox doubles); the one-pass kernel when there is none
  double* xs = nullptr;
  const size_t xs_bytes = (size_t)c->nj * ox * 8;
  const bool two_stage = det_two_stage(*c, oy);
  const bool captured = capture_active();
  if (two_stage && captured) {
    // the graph replays this detector: its scratch lives as long as the graph
    RSIM_CHECK(capture_alloc(xs_bytes, (void**)&xs));
  } else if (two_stage && hipMallocAsync((void**)&xs, xs_bytes, s) != hipSuccess) {
    (void)hipGetLastError();
    xs = nullptr;
  }
  const int rc = det_launch(u, nx, ny, dtype, dx, dy, ox, oy, z, weight, accumulate, out_dev, *c,
                            xs, s);
  if (xs && !captured) (void)hipFreeAsync(xs, s);
  return rc;
}

// ---------------------------------------------------------------------------
// The detector without allocations, for callers that cannot allocate device
// memory.  Its plan and its x-integrals live in one static device array of
// kNoAllocScratchBytes per GPU (held until the process ends); every call on a
// GPU reuses that GPU's array, so calls that overlap on one GPU race.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocScratchBytes = 64ull << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
std::mutex g_scratch_mu;
std::map<int, unsigned char*> g_scratch_ptr;   // device -> address of that device's array

// The array's address on the current device, looked up once per device (the
// array's name in host code is only a stand-in for the symbol).
int noalloc_scratch(unsigned char** out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  const int dev = current_device();
  auto it = g_scratch_ptr.find(dev);
  if (it != g_scratch_ptr.end()) {
    *out = it->second;
    return RSIM_OK;
  }
  void* p = nullptr;
  RSIM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) return fail(RSIM_E_HIP, "no device address for the detector scratch array");
  g_scratch_ptr[dev] = static_cast<unsigned char*>(p);
  *out = static_cast<unsigned char*>(p);
  return RSIM_OK;
}
}  // namespace

size_t detector_noalloc_capacity() { return kNoAllocScratchBytes; }

int detector_area_v1_noalloc(const void* u, int nx, int ny, int dtype, double dx, double dy,
                             int ox, int oy, double dsx, double dsy, double z, double weight,
                             double* out_dev, hipStream_t s) {
  RSIM_CHECK(check_detector(ox, oy, z));
  if (capture_active())
    return fail(RSIM_E_UNSUPPORTED, "the no-allocation detector uploads its plan synchronously "
                "and cannot be recorded into a graph; record rsim_detector_area_v1_accumulate");
  unsigned char* base = nullptr;
  RSIM_CHECK(noalloc_scratch(&base));
  DetPlanHost P;
  det_plan(nx, ny, dx, dy, ox, oy, dsx, dsy, &P);
  Blob b = make_blob({&P.xi0, &P.xi1, &P.jlo, &P.jhi},
                     {&P.xf0, &P.xf1, &P.xdet, &P.ylo, &P.yhi, &P.ydet});
  DetCached c;
  det_plan_bind(P, b, base, ox, oy, ny, &c);
  const size_t plan_bytes = (b.bytes.size() + 255) & ~size_t(255);
  const bool two = det_two_stage(c, oy);
  const size_t xs_bytes = two ? (size_t)c.nj * ox * 8 : 0;
  if (plan_bytes + xs_bytes > kNoAllocScratchBytes)
    return fail(RSIM_E_INVALID_ARG,
                "detector of %d x %d pixels over %d field rows needs %zu bytes of scratch; the "
                "no-allocation detector holds %zu (use rsim_detector_area_v1_accumulate)",
                ox, oy, c.nj, plan_bytes + xs_bytes, kNoAllocScratchBytes);
  // the plan replaces the last call's, after the work already on the stream
  // (which may still read it); the host copy must outlive the transfer
  RSIM_HIP(hipMemcpyAsync(base, b.bytes.data(), b.bytes.size(), hipMemcpyHostToDevice, s));
  RSIM_HIP(hipStreamSynchronize(s));
  return det_launch(u, nx, ny, dtype, dx, dy, ox, oy, z, weight, true, out_dev, c,
                    two ? reinterpret_cast<double*>(base + plan_bytes) : nullptr, s);
}

int detector_xbin_plan(int nx, int ny, double dx, double dy, int ox, int oy, double dsx,
                       double dsy, XBinArgs* xb, bool* ok) {
  *ok = false;
  const DetCached* c = nullptr;
  RSIM_CHECK(det_plan_dev(nx, ny, dx, dy, ox, oy, dsx, dsy, &c));
  if (!c->mono || c->nj <= 0 || oy > 65535) return RSIM_OK;
  *xb = XBinArgs{c->pl.xi0, c->pl.xi1, c->pl.xf0, c->pl.xf1, nullptr, ox, c->j0, c->nj, dx};
  *ok = true;
  return RSIM_OK;
}

int detector_ybin(const double* xs, int nx, int ny, double dx, double dy, int ox, int oy,
                  double dsx, double dsy, double z, double weight, bool accumulate,
                  double* out_dev, hipStream_t s) {
  RSIM_CHECK(check_detector(ox, oy, z));
  const DetCached* c = nullptr;
  RSIM_CHECK(det_plan_dev(nx, ny, dx, dy, ox, oy, dsx, dsy, &c));
  if (oy > 65535) return fail(RSIM_E_INVALID_ARG, "detector rows %d exceed 65535", oy);
  hipLaunchKernelGGL(area_v1_cols_kernel, dim3((ox + 255) / 256, oy), dim3(256), 0, s, xs, ox,
                     oy, c->j0, c->pl, dy, c->gyl, z, weight, accumulate ? 1 : 0, out_dev);
  RSIM_LAUNCHED();
  return RSIM_OK;
}

// ---------------------------------------------------------------------------
// Adjoint of area_v1.  The forward is I[q,p] = cos(q,p) * sum_{j,i} oy(q,j)
// wx(p,i) |v[j,i]|^2 with the column weight wx(p,i) = dx * ([i0 <= i < i1] +
// f1 [i == i1] - f0 [i == i0]) and the row overlap oy(q,j) of area_v1_kernel.
// Given gI = dL/dI (real), dL/dv* = v * G with G[j,i] = sum_{q,p} gI[q,p]
// cos(q,p) oy(q,j) wx(p,i).  Each field cell gathers from the detector
// pixels whose column / row ranges contain it: p in [plo[i], phi[i]), q in
// [qlo[j], qhi[j]) (both ranges are monotone in the cell index).
// ---------------------------------------------------------------------------
struct DetAdjDev {
  const int* plo;
  const int* phi;
  const int* qlo;
  const int* qhi;
};

template <typename R>
__global__ void __launch_bounds__(256) area_v1_adjoint_kernel(
    const cpx<R>* __restrict__ v, cpx<R>* __restrict__ g, int nx, int ny, int ox, DetPlanDev pl,
    DetAdjDev ap, double dx, double dy, double gyl, double z, const double* __restrict__ gimg) {
  const long long n = (long long)nx * ny;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < n;
       idx += (long long)gridDim.x * blockDim.x) {
    const int j = (int)(idx / nx);
    const int i = (int)(idx - (long long)j * nx);
    const double cell_lo = gyl + j * dy;
    const double cell_hi = cell_lo + dy;
    double G = 0.0;
    for (int q = ap.qlo[j]; q < ap.qhi[j]; ++q) {
      const double ov = fmin(cell_hi, pl.yhi[q]) - fmax(cell_lo, pl.ylo[q]);
      if (!(ov > 0.0)) continue;
      const double yd = pl.ydet[q];
      double h = 0.0;
      for (int p = ap.plo[i]; p < ap.phi[i]; ++p) {
        const int i0 = pl.xi0[p], i1 = pl.xi1[p];
        double w = (i >= i0 && i < i1) ? dx : 0.0;
        if (i == i1) w += pl.xf1[p] * dx;
        if (i == i0) w -= pl.xf0[p] * dx;
        const double xd = pl.xdet[p];
        const double cosine = z / sqrt(xd * xd + yd * yd + z * z);
        h += gimg[(long long)q * ox + p] * cosine * w;
      }
      G += ov * h;
    }
    const cpx<R> a = v[idx];
    g[idx] = mk((R)((double)a.x * G), (R)((double)a.y * G));
  }
}

int detector_area_v1_adjoint(const void* v, void* g, int nx, int ny, int dtype, double dx,
                             double dy, int ox, int oy, double dsx, double dsy, double z,
                             const double* gimg_dev, hipStream_t s) {
  RSIM_CHECK(check_detector(ox, oy, z));
  DetPlanHost P;
  det_plan(nx, ny, dx, dy, ox, oy, dsx, dsy, &P);
  // pixels containing column i: xi0[p] <= i <= xi1[p]; rows: jlo[q] <= j < jhi[q]
  std::vector<int> plo(nx), phi(nx), qlo(ny), qhi(ny);
  {
    int a = 0, b = 0;
    for (int i = 0; i < nx; ++i) {
      while (a < ox && P.xi1[a] < i) ++a;
      while (b < ox && P.xi0[b] <= i) ++b;
      plo[i] = a;
      phi[i] = std::max(a, b);
    }
    a = b = 0;
    for (int j = 0; j < ny; ++j) {
      while (a < oy && P.jhi[a] <= j) ++a;
      while (b < oy && P.jlo[b] <= j) ++b;
      qlo[j] = a;
      qhi[j] = std::max(a, b);
    }
  }
  Blob b = make_blob({&P.xi0, &P.xi1, &plo, &phi, &qlo, &qhi},
                     {&P.xf0, &P.xf1, &P.xdet, &P.ylo, &P.yhi, &P.ydet});
  unsigned char* d = nullptr;
  RSIM_HIP(hipMallocAsync((void**)&d, b.bytes.size(), s));
  RSIM_HIP(hipMemcpyAsync(d, b.bytes.data(), b.bytes.size(), hipMemcpyHostToDevice, s));
  DetPlanDev pl{};
  pl.xi0 = (const int*)(d + b.ioff[0]); pl.xi1 = (const int*)(d + b.ioff[1]);
  pl.xf0 = (const double*)(d + b.doff[0]); pl.xf1 = (const double*)(d + b.doff[1]);
  pl.xdet = (const double*)(d + b.doff[2]); pl.ylo = (const double*)(d + b.doff[3]);
  pl.yhi = (const double*)(d + b.doff[4]); pl.ydet = (const double*)(d + b.doff[5]);
  DetAdjDev ap;
  ap.plo = (const int*)(d + b.ioff[2]); ap.phi = (const int*)(d + b.ioff[3]);
  ap.qlo = (const int*)(d + b.ioff[4]); ap.qhi = (const int*)(d + b.ioff[5]);
  const long long n = (long long)nx * ny;
  const int blocks = (int)std::min<long long>((n + 255) / 256, 65536);
  if (dtype == RSIM_C64)
    hipLaunchKernelGGL(area_v1_adjoint_kernel<float>, dim3(blocks), dim3(256), 0, s,
                       (const cpx<float>*)v, (cpx<float>*)g, nx, ny, ox, pl, ap, dx, dy, P.gyl, z,
                       gimg_dev);
  else
    hipLaunchKernelGGL(area_v1_adjoint_kernel<double>, dim3(blocks), dim3(256), 0, s,
                       (const cpx<double>*)v, (cpx<double>*)g, nx, ny, ox, pl, ap, dx, dy, P.gyl,
                       z, gimg_dev);
  RSIM_LAUNCHED();
  RSIM_HIP(hipFreeAsync(d, s));
  RSIM_HIP(hipStreamSynchronize(s));
  return RSIM_OK;
}

// ---------------------------------------------------------------------------
// Volume painting: seeded synthetic samples (spheres, later ones overwrite)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) paint_sphere_kernel(uint8_t* ids, int nz, int ny, int nx,
                                                           int z0, int y0, int x0, int bz, int by,
                                                           int bx, double cz, double cy,
                                                           double cx, double r2, uint8_t mat) {
  const long long n = (long long)bz * by * bx;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < n;
       idx += (long long)gridDim.x * blockDim.x) {
    const int lx = (int)(idx % bx);
    const int ly = (int)((idx / bx) % by);
    const int lz = (int)(idx / ((long long)bx * by));
    const int z = z0 + lz, y = y0 + ly, x = x0 + lx;
    const double dz = z + 0.5 - cz, dy = y + 0.5 - cy, dx = x + 0.5 - cx;
    if (dz * dz + dy * dy + dx * dx <= r2) ids[((long long)z * ny + y) * nx + x] = mat;
  }
}

int paint_spheres(rsim_volume* vol, const double* sph, const int* mats, int n, hipStream_t s) {
  for (int k = 0; k < n; ++k) {
    const double cz = sph[4 * k], cy = sph[4 * k + 1], cx = sph[4 * k + 2], r = sph[4 * k + 3];
    if (mats[k] < 0 || mats[k] >= vol->n_db || mats[k] > 255)
      return fail(RSIM_E_INVALID_ARG, "sphere %d: material id %d out of range", k, mats[k]);
    if (!(r > 0)) continue;
    const int z0 = std::max(0, (int)std::floor(cz - r - 1)), z1 = std::min(vol->nz, (int)std::ceil(cz + r + 1));
    const int y0 = std::max(0, (int)std::floor(cy - r - 1)), y1 = std::min(vol->ny_g, (int)std::ceil(cy + r + 1));
    const int x0 = std::max(0, (int)std::floor(cx - r - 1)), x1 = std::min(vol->nx_g, (int)std::ceil(cx + r + 1));
    if (z1 <= z0 || y1 <= y0 || x1 <= x0) continue;
    const long long cnt = (long long)(z1 - z0) * (y1 - y0) * (x1 - x0);
    int blocks = (int)std::min<long long>((cnt + 255) / 256, 65536);
    hipLaunchKernelGGL(paint_sphere_kernel, dim3(blocks), dim3(256), 0, s, vol->ids, vol->nz,
                       vol->ny_g, vol->nx_g, z0, y0, x0, z1 - z0, y1 - y0, x1 - x0, cz, cy, cx,
                       r * r, (uint8_t)mats[k]);
    RSIM_LAUNCHED();
  }
  return RSIM_OK;
}

// ---------------------------------------------------------------------------
// Zero fill as a kernel (a graph replays it like any other launch; graphs
// here hold no memset nodes)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) zero_fill_kernel(double2* p, long long n2, double* tail,
                                                        int n_tail) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n2; i += stride)
    p[i] = make_double2(0.0, 0.0);
  if (blockIdx.x == 0 && (int)threadIdx.x < n_tail) tail[threadIdx.x] = 0.0;
}

int fill_zero(void* p, size_t bytes, hipStream_t s) {
  if (bytes == 0) return RSIM_OK;
  if ((reinterpret_cast<uintptr_t>(p) & 15) || (bytes & 7))
    return fail(RSIM_E_INVALID_ARG, "zero fill needs a 16-byte aligned buffer of 8-byte words");
  const long long n2 = (long long)(bytes / 16);
  const int n_tail = (int)((bytes % 16) / 8);
  const int blocks = (int)std::max<long long>(1, std::min<long long>((n2 + 255) / 256, 8192));
  hipLaunchKernelGGL(zero_fill_kernel, dim3(blocks), dim3(256), 0, s, (double2*)p, n2,
                     (double*)((char*)p + 16 * n2), n_tail);
  RSIM_LAUNCHED();
  return RSIM_OK;
}

// ---------------------------------------------------------------------------
// Device copy (the measured copy ceiling beside hipMemcpy): each workgroup
// copies one contiguous 128 KiB tile, 16 loads of 8 bytes per thread in
// flight before its 16 stores; every wavefront access is 512 contiguous
// bytes.
// ---------------------------------------------------------------------------
constexpr int kCopyThreads = 1024, kCopyPer = 16;
__global__ void __launch_bounds__(kCopyThreads) stream_copy_kernel(const uint2* __restrict__ src,
                                                                   uint2* __restrict__ dst,
                                                                   long long n) {
  const long long base = (long long)blockIdx.x * (kCopyThreads * kCopyPer) + threadIdx.x;
  uint2 v[kCopyPer];
  if (base + (long long)(kCopyPer - 1) * kCopyThreads < n) {
#pragma unroll
    for (int e = 0; e < kCopyPer; ++e) v[e] = src[base + e * kCopyThreads];
#pragma unroll
    for (int e = 0; e < kCopyPer; ++e) dst[base + e * kCopyThreads] = v[e];
  } else {
    for (int e = 0; e < kCopyPer; ++e)
      if (base + e * kCopyThreads < n) dst[base + e * kCopyThreads] = src[base + e * kCopyThreads];
  }
}

int stream_copy(void* dst, const void* src, size_t bytes, hipStream_t s) {
  if ((reinterpret_cast<uintptr_t>(dst) | reinterpret_cast<uintptr_t>(src) | bytes) & 7)
    return fail(RSIM_E_INVALID_ARG, "stream copy needs 8-byte aligned buffers and sizes");
  const long long n = (long long)(bytes / 8);
  const long long tile = (long long)kCopyThreads * kCopyPer;
  const long long blocks = (n + tile - 1) / tile;
  if (blocks > 0x7fffffffLL) return fail(RSIM_E_INVALID_ARG, "stream copy of %zu bytes is too large", bytes);
  hipLaunchKernelGGL(stream_copy_kernel, dim3((unsigned)blocks), dim3(kCopyThreads), 0, s,
                     (const uint2*)src, (uint2*)dst, n);
  RSIM_LAUNCHED();
  return RSIM_OK;
}

}  // namespace rsim
