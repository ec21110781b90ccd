// Launcher of the plain register-resident FFT kernels for one precision.
// Included by rsim_fft_reg_c64.hip / rsim_fft_reg_c128.hip.
#pragma once

#include "rsim_fft_cfg.hpp"
#include "rsim_internal.hpp"

#include <type_traits>

namespace rsim {

// complex64 transforms use compensated FP32 butterflies (twc tables): plain
// FP32 roundings of the radix-16 constants and twiddles bias |W| the same way
// in every transform, which accumulates over a multislice run.
template <typename R>
using RegTw = typename std::conditional<std::is_same<R, float>::value, twc, R>::type;

template <typename R, int N, int DIR, bool COL>
static void launch_one(const void* in, void* out, long long es, long long ss, long long bs,
                       int nseq, int batch, const void* const* tw, double scale, hipStream_t s) {
  using F = typename Sched<N>::F;
  using TW = RegTw<R>;
  StageTw<TW> t;
  using TP = typename std::remove_reference<decltype(t.t[0])>::type;
  for (int i = 0; i < 4; ++i) t.t[i] = reinterpret_cast<TP>(tw[i]);
  const long long esz = sizeof(cpx<R>);
  const uint32_t es_b = (uint32_t)(es * esz), ss_b = (uint32_t)(ss * esz);
  const uint32_t nbytes = (uint32_t)(((long long)(nseq - 1) * ss + (long long)(N - 1) * es + 1) * esz);
  using S = typename std::conditional<COL, ColShape<R, N>, RowShape<R, N>>::type;
  dim3 grid((nseq + S::C - 1) / S::C, batch);
  hipLaunchKernelGGL((fft_reg_kernel<R, F, DIR, S::C, S::PS, S::SPLIT, TW>), grid, dim3(F::TC * S::C), 0,
                     s, reinterpret_cast<const cpx<R>*>(in), reinterpret_cast<cpx<R>*>(out), es_b,
                     ss_b, bs, nbytes, nseq, t, (R)scale);
}

// Byte extents of one batch must fit the 32-bit buffer offsets.
template <typename R, int N>
static bool fits32(long long es, long long ss, int nseq) {
  const long long ext = ((long long)(nseq - 1) * ss + (long long)(N - 1) * es + 1) * (long long)sizeof(cpx<R>);
  return ext < (1LL << 32) && es * (long long)sizeof(cpx<R>) < (1LL << 32) &&
         ss * (long long)sizeof(cpx<R>) < (1LL << 32);
}

template <typename R, int N>
static bool launch_n(int dir, bool col, const void* in, void* out, long long es, long long ss,
                     long long bs, int nseq, int batch, const void* const* tw, double scale,
                     hipStream_t s) {
  if (!fits32<R, N>(es, ss, nseq)) return false;
  if (dir < 0) {
    if (col) launch_one<R, N, -1, true>(in, out, es, ss, bs, nseq, batch, tw, scale, s);
    else launch_one<R, N, -1, false>(in, out, es, ss, bs, nseq, batch, tw, scale, s);
  } else {
    if (col) launch_one<R, N, 1, true>(in, out, es, ss, bs, nseq, batch, tw, scale, s);
    else launch_one<R, N, 1, false>(in, out, es, ss, bs, nseq, batch, tw, scale, s);
  }
  return true;
}

template <typename R>
static bool launch_reg(int dir, int N, bool col, const void* in, void* out, long long es,
                       long long ss, long long bs, int nseq, int batch, const void* const* tw,
                       double scale, hipStream_t s) {
  switch (N) {
    case 256: return launch_n<R, 256>(dir, col, in, out, es, ss, bs, nseq, batch, tw, scale, s);
    case 512: return launch_n<R, 512>(dir, col, in, out, es, ss, bs, nseq, batch, tw, scale, s);
    case 1024: return launch_n<R, 1024>(dir, col, in, out, es, ss, bs, nseq, batch, tw, scale, s);
    case 2048: return launch_n<R, 2048>(dir, col, in, out, es, ss, bs, nseq, batch, tw, scale, s);
    case 4096: return launch_n<R, 4096>(dir, col, in, out, es, ss, bs, nseq, batch, tw, scale, s);
    case 8192: return launch_n<R, 8192>(dir, col, in, out, es, ss, bs, nseq, batch, tw, scale, s);
    default: return false;
  }
}

}  // namespace rsim
