// Differentiable multislice through per-voxel delta/beta volumes.
//
// Volume: float32 delta and beta, (nz, ny, nx) C-order on the field grid (one
// voxel per field pixel), the direct per-voxel form of fast-wave's complex
// delta-beta grid (fast-wave/fwcuda/kernels.cu:411-491).  Slice s applies
// T_s = exp(-K (beta_s + i delta_s)), K = 2 pi pz / wl, which is big-wave's
// material_factor exp(-2 pi i pz/wl conj(delta + i beta))
// (big-wave/optical_element.py:39-42), then propagate_2d over pz
// (big-wave/propagation.py:396-442):  u_{s+1} = P (T_s u_s).
//
// Adjoint (Wirtinger calculus, L real): with g = dL/du* after a slice,
//   h   = P^H g        (P^H = propagation over -pz: conjugate chirp, same mask)
//   w   = T_s u_s
//   dL/d delta_s = 2 K Im(conj(h) w),  dL/d beta_s = -2 K Re(conj(h) w)
//   g  <- conj(T_s) h  (dL/du_s*)
// Forward fields are rebuilt from checkpoints stored every `every` slices
// (sqrt(n) checkpointing: ceil(n/every) stored fields plus every-1 rebuilt).
// This file is the plain engine (any size, complex64 or complex128); the
// fused complex64 engine's form of both passes is in rsim_multislice.hip.
#include <algorithm>
#include <cmath>

#include "rsim_device.hpp"
#include "rsim_internal.hpp"

namespace rsim {

int propagate(void* u, void* U, int nx, int ny, int dtype, double dx, double dy, double wl,
              double dz, double freq, bool skip_fft, hipStream_t s);

namespace {

__device__ __forceinline__ void tfac(float K, float d, float b, float* tr, float* ti) {
  float sn, cs;
  sincosf(K * d, &sn, &cs);
  const float a = expf(-K * b);
  *tr = a * cs;
  *ti = -a * sn;
}
__device__ __forceinline__ void tfac(double K, float d, float b, double* tr, double* ti) {
  double sn, cs;
  sincos(K * (double)d, &sn, &cs);
  const double a = exp(-K * (double)b);
  *tr = a * cs;
  *ti = -a * sn;
}

template <typename R>
__global__ void __launch_bounds__(256) transmission_db_kernel(cpx<R>* __restrict__ u, long long n,
                                                              const float* __restrict__ delta,
                                                              const float* __restrict__ beta,
                                                              R K) {
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < n;
       idx += (long long)gridDim.x * blockDim.x) {
    R tr, ti;
    tfac(K, delta[idx], beta[idx], &tr, &ti);
    const cpx<R> v = u[idx];
    u[idx] = mk(v.x * tr - v.y * ti, v.x * ti + v.y * tr);
  }
}

template <typename R>
__global__ void __launch_bounds__(256) adjoint_slice_kernel(cpx<R>* __restrict__ g,
                                                            const cpx<R>* __restrict__ u,
                                                            long long n,
                                                            const float* __restrict__ delta,
                                                            const float* __restrict__ beta, R K,
                                                            float* __restrict__ gd,
                                                            float* __restrict__ gb) {
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < n;
       idx += (long long)gridDim.x * blockDim.x) {
    R tr, ti;
    tfac(K, delta[idx], beta[idx], &tr, &ti);
    const cpx<R> v = u[idx];
    const cpx<R> h = g[idx];
    const R wr = v.x * tr - v.y * ti, wi = v.x * ti + v.y * tr;
    // conj(h) * w
    const R cr = h.x * wr + h.y * wi;
    const R ci = h.x * wi - h.y * wr;
    gd[idx] = (float)(R(2) * K * ci);
    gb[idx] = (float)(R(-2) * K * cr);
    g[idx] = mk(tr * h.x + ti * h.y, tr * h.y - ti * h.x);
  }
}

int launch_blocks(long long n) { return (int)std::min<long long>((n + 255) / 256, 65536); }

int transmission_db(void* u, int nx, int ny, int dtype, const float* delta, const float* beta,
                    double K, hipStream_t s) {
  const long long n = (long long)nx * ny;
  ktime_begin(KT_PW, s);
  if (dtype == RSIM_C64)
    hipLaunchKernelGGL(transmission_db_kernel<float>, dim3(launch_blocks(n)), dim3(256), 0, s,
                       (cpx<float>*)u, n, delta, beta, (float)K);
  else
    hipLaunchKernelGGL(transmission_db_kernel<double>, dim3(launch_blocks(n)), dim3(256), 0, s,
                       (cpx<double>*)u, n, delta, beta, K);
  ktime_end(KT_PW, s);
  RSIM_LAUNCHED();
  return RSIM_OK;
}

}  // namespace

int transmission_db_2d(void* u, int nx, int ny, int dtype, double wl, double pz,
                       const float* delta, const float* beta, hipStream_t s) {
  return transmission_db(u, nx, ny, dtype, delta, beta, ((2.0 * M_PI) * pz) / wl, s);
}

int multislice_db(void* u, void* U, int nx, int ny, int dtype, double dx, double dy, double wl,
                  double freq, const float* delta, const float* beta, double pz, int n,
                  void* ckpt, int every, int engine, int flags, hipStream_t s) {
  if (engine == RSIM_FUSED && fused_db_supported(nx, ny, dtype))
    return fused_multislice_db(u, U, nx, ny, dx, dy, wl, freq, delta, beta, pz, n, ckpt, every,
                               flags, s);
  const size_t fb = (size_t)nx * ny * elem_bytes(dtype);
  const size_t vs = (size_t)nx * ny;
  const double K = ((2.0 * M_PI) * pz) / wl;
  for (int k = 0; k < n; ++k) {
    if (cancel_requested()) return fail(RSIM_E_CANCELLED, "multislice cancelled at slice %d", k);
    if (ckpt && k % every == 0)
      RSIM_HIP(hipMemcpyAsync((char*)ckpt + (size_t)(k / every) * fb, u, fb,
                              hipMemcpyDeviceToDevice, s));
    RSIM_CHECK(transmission_db(u, nx, ny, dtype, delta + k * vs, beta + k * vs, K, s));
    RSIM_CHECK(propagate(u, U, nx, ny, dtype, dx, dy, wl, pz, freq, false, s));
    if (progress_due(k + 1, n)) {
      RSIM_HIP(hipStreamSynchronize(s));
      report_progress("multislice", k + 1, n);
    }
  }
  return RSIM_OK;
}

int multislice_db_adjoint(void* g, void* U, int nx, int ny, int dtype, double dx, double dy,
                          double wl, double freq, const float* delta, const float* beta,
                          double pz, int n, const void* ckpt, int every, void* work,
                          float* gdelta, float* gbeta, int engine, int flags, hipStream_t s) {
  if (engine == RSIM_FUSED && fused_db_supported(nx, ny, dtype))
    return fused_multislice_db_adjoint(g, U, nx, ny, dx, dy, wl, freq, delta, beta, pz, n, ckpt,
                                       every, work, gdelta, gbeta, flags, s);
  const size_t fb = (size_t)nx * ny * elem_bytes(dtype);
  const size_t vs = (size_t)nx * ny;
  const long long np = (long long)vs;
  const double K = ((2.0 * M_PI) * pz) / wl;
  const int nseg = (n + every - 1) / every;
  long long done = 0;
  for (int seg = nseg - 1; seg >= 0; --seg) {
    const int s0 = seg * every;
    const int s1 = std::min(n, s0 + every);
    // field before slice k of this segment: the checkpoint, then the rebuilt fields
    auto field = [&](int k) -> void* {
      return k == s0 ? (void*)((const char*)ckpt + (size_t)seg * fb)
                     : (void*)((char*)work + (size_t)(k - s0 - 1) * fb);
    };
    for (int k = s0 + 1; k < s1; ++k) {
      void* dst = field(k);
      RSIM_HIP(hipMemcpyAsync(dst, field(k - 1), fb, hipMemcpyDeviceToDevice, s));
      RSIM_CHECK(transmission_db(dst, nx, ny, dtype, delta + (k - 1) * vs, beta + (k - 1) * vs, K, s));
      RSIM_CHECK(propagate(dst, U, nx, ny, dtype, dx, dy, wl, pz, freq, false, s));
    }
    for (int k = s1 - 1; k >= s0; --k) {
      if (cancel_requested()) return fail(RSIM_E_CANCELLED, "adjoint cancelled at slice %d", k);
      RSIM_CHECK(propagate(g, U, nx, ny, dtype, dx, dy, wl, -pz, freq, false, s));
      ktime_begin(KT_PW, s);
      if (dtype == RSIM_C64)
        hipLaunchKernelGGL(adjoint_slice_kernel<float>, dim3(launch_blocks(np)), dim3(256), 0, s,
                           (cpx<float>*)g, (const cpx<float>*)field(k), np, delta + k * vs,
                           beta + k * vs, (float)K, gdelta + k * vs, gbeta + k * vs);
      else
        hipLaunchKernelGGL(adjoint_slice_kernel<double>, dim3(launch_blocks(np)), dim3(256), 0,
                           s, (cpx<double>*)g, (const cpx<double>*)field(k), np, delta + k * vs,
                           beta + k * vs, K, gdelta + k * vs, gbeta + k * vs);
      ktime_end(KT_PW, s);
      RSIM_LAUNCHED();
      ++done;
      if (progress_due(done, n)) {
        RSIM_HIP(hipStreamSynchronize(s));
        report_progress("multislice_adjoint", done, n);
      }
    }
  }
  return RSIM_OK;
}

}  // namespace rsim
