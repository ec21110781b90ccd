// Launcher of the plain register-resident FFT kernels for one precision.
// Included by rsim_fft_reg_c64.hip / rsim_fft_reg_c128.hip.
#pragma once

#include "rsim_fft_cfg.hpp"
#include "rsim_internal.hpp"

namespace rsim {

template <typename R, int N, int DIR, bool COL>
static void launch_one(const void* in, void* out, long long es, long long ss, long long bs,
                       int nseq, int batch, const void* const* tw, double scale, hipStream_t s) {
  using F = typename Sched<N>::F;
  StageTw<R> t;
  for (int i = 0; i < 4; ++i) t.t[i] = reinterpret_cast<const cpx<R>*>(tw[i]);
  if constexpr (COL) {
    using S = ColShape<R, N>;
    dim3 grid((nseq + S::C - 1) / S::C, batch);
    hipLaunchKernelGGL((fft_reg_kernel<R, F, DIR, S::C, S::PS, S::SPLIT>), grid, dim3(F::TC * S::C),
                       0, s, reinterpret_cast<const cpx<R>*>(in), reinterpret_cast<cpx<R>*>(out),
                       es, ss, bs, nseq, t, (R)scale);
  } else {
    using S = RowShape<R, N>;
    dim3 grid((nseq + S::C - 1) / S::C, batch);
    hipLaunchKernelGGL((fft_reg_kernel<R, F, DIR, S::C, S::PS, S::SPLIT>), grid, dim3(F::TC * S::C),
                       0, s, reinterpret_cast<const cpx<R>*>(in), reinterpret_cast<cpx<R>*>(out),
                       es, ss, bs, nseq, t, (R)scale);
  }
}

template <typename R, int N>
static void launch_n(int dir, bool col, const void* in, void* out, long long es, long long ss,
                     long long bs, int nseq, int batch, const void* const* tw, double scale,
                     hipStream_t s) {
  if (dir < 0) {
    if (col) launch_one<R, N, -1, true>(in, out, es, ss, bs, nseq, batch, tw, scale, s);
    else launch_one<R, N, -1, false>(in, out, es, ss, bs, nseq, batch, tw, scale, s);
  } else {
    if (col) launch_one<R, N, 1, true>(in, out, es, ss, bs, nseq, batch, tw, scale, s);
    else launch_one<R, N, 1, false>(in, out, es, ss, bs, nseq, batch, tw, scale, s);
  }
}

template <typename R>
static bool launch_reg(int dir, int N, bool col, const void* in, void* out, long long es,
                       long long ss, long long bs, int nseq, int batch, const void* const* tw,
                       double scale, hipStream_t s) {
  switch (N) {
    case 256: launch_n<R, 256>(dir, col, in, out, es, ss, bs, nseq, batch, tw, scale, s); return true;
    case 512: launch_n<R, 512>(dir, col, in, out, es, ss, bs, nseq, batch, tw, scale, s); return true;
    case 1024: launch_n<R, 1024>(dir, col, in, out, es, ss, bs, nseq, batch, tw, scale, s); return true;
    case 2048: launch_n<R, 2048>(dir, col, in, out, es, ss, bs, nseq, batch, tw, scale, s); return true;
    case 4096: launch_n<R, 4096>(dir, col, in, out, es, ss, bs, nseq, batch, tw, scale, s); return true;
    case 8192: launch_n<R, 8192>(dir, col, in, out, es, ss, bs, nseq, batch, tw, scale, s); return true;
    default: return false;
  }
}

}  // namespace rsim
