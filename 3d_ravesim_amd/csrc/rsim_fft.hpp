// Stockham autosort FFT kernels for gfx950.
//
// Two kernel families share one stage formula.  For a length-N transform
// factored N = R_0 * R_1 * ... and stage s with radix R and stride
// L = R_0*...*R_{s-1}, butterfly b in [0, N/R) with k = b mod L reads
//   x_j = in[b + j*N/R] * W_{L*R}^{j*k}        (j in [0, R))
// applies a length-R DFT and writes
//   out[(b - k)*R + k + m*L] = y_m              (m in [0, R)).
// After the last stage the data are in natural order.
//
// * fft_lds_kernel: runtime factorization (radix 2,3,4,5,8,16), the whole
//   tile of C sequences staged in LDS with ping-pong buffers.  Handles every
//   supported size up to the LDS capacity; used for small and uncommon sizes.
// * reg_fft (fft_reg_kernel and the fused multislice kernels): compile-time
//   factorization, every thread keeps E = N/TC elements in VGPRs, LDS is only
//   the exchange buffer between stages (optionally re/im split to halve it).
//   The first stage loads from HBM and the last stage stores to HBM directly,
//   and pointwise work (transmission, chirp, scaling) is applied in registers.
#pragma once

#include <type_traits>

#include "rsim_device.hpp"

namespace rsim {

constexpr int kMaxStages = 12;

struct AxisDesc {
  int N;
  int nst;
  int radix[kMaxStages];
  int ls[kMaxStages];
  long long es_in, es_out;   // element stride along the transform
  long long ss_in, ss_out;   // stride between neighbouring sequences
  long long bs_in, bs_out;   // stride between batches (grid.y)
  int nseq;                  // sequences per batch
  int C;                     // sequences per workgroup tile
  double scale;              // output multiplier (1 or 1/N)
  const void* tw;            // W_N^m, m in [0, N), forward sign, in the transform's precision
  // Optional post-twiddle (a four-step transform's middle factor): output
  // element pos of a sequence is multiplied by W_P^(i*pos) (conjugated for the
  // inverse), i = the sequence index (ptw_mode 1) or the batch index (2).
  int ptw_mode;
  long long ptw_P;
  const void* ptw;           // W_P^m, m in [0, P)
};

template <int RR, int DIR, typename R>
__device__ __forceinline__ void lds_bfly(const cpx<R>* __restrict__ src, cpx<R>* __restrict__ dst,
                                         const cpx<R>* __restrict__ tw, int N, int LS, int C, int c,
                                         int b) {
  const int nb = N / RR;
  const int k = b % LS;
  cpx<R> x[RR];
#pragma unroll
  for (int j = 0; j < RR; ++j) x[j] = src[(b + j * nb) * C + c];
  if (LS > 1) {
    const int step = N / (LS * RR);
#pragma unroll
    for (int j = 1; j < RR; ++j) x[j] = twmul<DIR>(x[j], tw[j * k * step]);
  }
  dft<RR, DIR>(x);
  const int base = (b - k) * RR + k;
#pragma unroll
  for (int m = 0; m < RR; ++m) dst[(base + m * LS) * C + c] = x[m];
}

template <typename R, int DIR>
__global__ void __launch_bounds__(256) fft_lds_kernel(const cpx<R>* __restrict__ in,
                                                      cpx<R>* __restrict__ out, AxisDesc d) {
  extern __shared__ unsigned char smem_raw[];
  cpx<R>* buf0 = reinterpret_cast<cpx<R>*>(smem_raw);
  cpx<R>* buf1 = buf0 + (size_t)d.C * d.N;
  const cpx<R>* tw = reinterpret_cast<const cpx<R>*>(d.tw);
  const int T = blockDim.x;
  const int tid = threadIdx.x;
  const int N = d.N, C = d.C;
  const int seq0 = blockIdx.x * C;
  const long long boff_in = (long long)blockIdx.y * d.bs_in;
  const long long boff_out = (long long)blockIdx.y * d.bs_out;
  const bool pos_fast_in = d.es_in <= d.ss_in;
  const bool pos_fast_out = d.es_out <= d.ss_out;
  const int total = C * N;

  for (int idx = tid; idx < total; idx += T) {
    int pos, c;
    if (pos_fast_in) { pos = idx % N; c = idx / N; }
    else { c = idx % C; pos = idx / C; }
    const int seq = seq0 + c;
    cpx<R> v = mk(R(0), R(0));
    if (seq < d.nseq) v = in[boff_in + (long long)seq * d.ss_in + (long long)pos * d.es_in];
    buf0[pos * C + c] = v;
  }
  __syncthreads();

  cpx<R>* src = buf0;
  cpx<R>* dst = buf1;
  for (int s = 0; s < d.nst; ++s) {
    const int RR = d.radix[s];
    const int LS = d.ls[s];
    const int work = C * (N / RR);
    for (int w = tid; w < work; w += T) {
      const int c = w % C;
      const int b = w / C;
      switch (RR) {
        case 2: lds_bfly<2, DIR>(src, dst, tw, N, LS, C, c, b); break;
        case 3: lds_bfly<3, DIR>(src, dst, tw, N, LS, C, c, b); break;
        case 4: lds_bfly<4, DIR>(src, dst, tw, N, LS, C, c, b); break;
        case 5: lds_bfly<5, DIR>(src, dst, tw, N, LS, C, c, b); break;
        case 8: lds_bfly<8, DIR>(src, dst, tw, N, LS, C, c, b); break;
        default: lds_bfly<16, DIR>(src, dst, tw, N, LS, C, c, b); break;
      }
    }
    __syncthreads();
    cpx<R>* t = src;
    src = dst;
    dst = t;
  }

  const R sc = R(d.scale);
  for (int idx = tid; idx < total; idx += T) {
    int pos, c;
    if (pos_fast_out) { pos = idx % N; c = idx / N; }
    else { c = idx % C; pos = idx / C; }
    const int seq = seq0 + c;
    if (seq < d.nseq) {
      cpx<R> v = src[pos * C + c];
      if (d.ptw_mode) {
        const long long i = d.ptw_mode == 1 ? seq : (long long)blockIdx.y;
        v = twmul<DIR>(v, reinterpret_cast<const cpx<R>*>(d.ptw)[i * pos]);
      }
      if (d.scale != 1.0) v = cscale(v, sc);
      out[boff_out + (long long)seq * d.ss_out + (long long)pos * d.es_out] = v;
    }
  }
}

// ---------------------------------------------------------------------------
// Register-resident FFT (compile-time factorization)
// ---------------------------------------------------------------------------

// Factorization descriptor: N = R0*R1*R2*R3 (trailing 1s unused), E elements
// per thread, TC = N / E threads per sequence.  E must be a multiple of
// every radix.
template <int N_, int E_, int R0_, int R1_ = 1, int R2_ = 1, int R3_ = 1>
struct RegFac {
  static constexpr int N = N_;
  static constexpr int E = E_;
  static constexpr int TC = N_ / E_;
  static constexpr int R0 = R0_, R1 = R1_, R2 = R2_, R3 = R3_;
  static constexpr int NS = 1 + (R1_ > 1) + (R2_ > 1) + (R3_ > 1);
  static_assert(R0_ * R1_ * R2_ * R3_ == N_, "radices must multiply to N");
  static_assert(E_ % R0_ == 0 && E_ % R1_ == 0 && E_ % R2_ == 0 && E_ % R3_ == 0,
                "E must be a multiple of every radix");
  static constexpr int radix(int s) { return s == 0 ? R0 : s == 1 ? R1 : s == 2 ? R2 : R3; }
  static constexpr int stride(int s) {
    return s == 0 ? 1 : s == 1 ? R0 : s == 2 ? R0 * R1 : R0 * R1 * R2;
  }
  static constexpr int last_radix() { return radix(NS - 1); }
};

// Twiddle tables for a RegFac: table s (s >= 1) holds, at [k*R_s + j],
// W_{L_s R_s}^(j k) for k in [0, L_s), j in [0, R_s), forward sign -- the R_s
// twiddles of one butterfly are contiguous, so their addresses are one base
// plus compile-time offsets.
template <typename R>
struct StageTw {
  const cpx<R>* t[4];
};
// compensated single-precision tables: FP32 hi + lo per twiddle
template <>
struct StageTw<twc> {
  const twc* t[4];
};

// LDS padding code: PS = (shift << 8) | k inserts k pad slots after every
// 2^shift slots (PS = 0: no padding).  Chosen per exchange pattern to spread
// the strided stage writes over the banks.
constexpr int pad_code(int shift, int k) { return (shift << 8) | k; }
template <int PS>
__device__ __forceinline__ int lpad(int e) {
  if constexpr (PS == 0) return e;
  else return e + (PS & 255) * (e >> (PS >> 8));
}

template <class F, int C, int PS>
struct LdsSize {
  static constexpr int slots =
      F::N * C + ((PS == 0) ? 0 : (PS & 255) * ((F::N * C) >> (PS >> 8))) + 1;
};

// Output position of element (i, m) of stage s (radix RR, stride LS).
template <class F, int RR, int LS>
__device__ __forceinline__ int out_pos(int t, int i, int m) {
  const int b = t + i * F::TC;
  const int k = b % LS;
  return (b - k) * RR + k + m * LS;
}
template <class F, int RR>
__device__ __forceinline__ int in_pos(int t, int i, int j) {
  return t + i * F::TC + j * (F::N / RR);
}

// LDS slot of position pos of sequence c: with PS = pad_code(shift, k), C a
// power of two and k a multiple of C, slot = lpad(pos * C + c) =
// (pos + q * (pos >> s)) * C + c with s = shift - log2(C), q = k / C.
template <int PS, int C>
struct PadPos {
  static constexpr int log2c =
      C >= 64 ? 6 : C >= 32 ? 5 : C >= 16 ? 4 : C >= 8 ? 3 : C >= 4 ? 2 : C >= 2 ? 1 : 0;
  static constexpr bool ok = (C & (C - 1)) == 0 && C <= 64 &&
                             (PS == 0 || ((PS & 255) % C == 0 && (PS >> 8) >= log2c));
  static constexpr int s = PS == 0 ? 0 : (PS >> 8) - log2c;
  static constexpr int q = PS == 0 ? 0 : (PS & 255) / C;
  // slot offset of position offset K whose low s bits add to a base without carry
  static constexpr int off(int K) { return (PS == 0 ? K : K + q * (K >> s)) * C; }
  __device__ static __forceinline__ int base(int pos, int c) {
    return (PS == 0 ? pos : pos + q * (pos >> s)) * C + c;
  }
};
// Whether the exchange into stage S has slot addresses of the form
// base(thread) + constant(i, m) on both sides (every output position of
// stage S-1 is pb(t) + i*TC*RP + m*LP with pb(t) = (t - t%LP)*RP + t%LP when
// TC is a multiple of LP; every input position of stage S is t + i*TC +
// j*N/RN), and the padding adds no carry between the two parts.
template <class F, int S, int C, int PS>
struct ExLinear {
  static constexpr int RP = F::radix(S - 1);
  static constexpr int LP = F::stride(S - 1);
  static constexpr int RN = F::radix(S);
  using P = PadPos<PS, C>;
  static constexpr int M = 1 << P::s;
  static constexpr bool wr = P::ok && F::TC % LP == 0 &&
                             (PS == 0 || ((F::TC * RP) % M == 0 &&
                                          (LP % M == 0 || (M % LP == 0 && (LP * RP) % M == 0))));
  static constexpr bool rd = P::ok && (PS == 0 || (F::TC % M == 0 && (F::N / RN) % M == 0));
};

// Exchange between stage S-1 (outputs in v) and stage S (inputs into v).
template <class F, int S, int C, int PS, bool SPLIT, typename R>
__device__ __forceinline__ void reg_exchange(cpx<R> (&v)[F::E], void* lds_raw, int t, int c) {
  constexpr int RP = F::radix(S - 1);
  constexpr int LP = F::stride(S - 1);
  constexpr int RN = F::radix(S);
  using X = ExLinear<F, S, C, PS>;
  using P = PadPos<PS, C>;
  // slot of output (i, m) of stage S-1 and of input (i, j) of stage S
  int wb = 0, rb = 0;
  if constexpr (X::wr) wb = P::base((t - t % LP) * RP + t % LP, c);
  if constexpr (X::rd) rb = P::base(t, c);
  auto wslot = [&](int i, int m) {
    if constexpr (X::wr) return wb + P::off(i * F::TC * RP + m * LP);
    else return lpad<PS>(out_pos<F, RP, LP>(t, i, m) * C + c);
  };
  auto rslot = [&](int i, int j) {
    if constexpr (X::rd) return rb + P::off(i * F::TC + j * (F::N / RN));
    else return lpad<PS>(in_pos<F, RN>(t, i, j) * C + c);
  };
  if constexpr (!SPLIT) {
    cpx<R>* lds = reinterpret_cast<cpx<R>*>(lds_raw);
#pragma unroll
    for (int i = 0; i < F::E / RP; ++i)
#pragma unroll
      for (int m = 0; m < RP; ++m) lds[wslot(i, m)] = v[i * RP + m];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < F::E / RN; ++i)
#pragma unroll
      for (int j = 0; j < RN; ++j) v[i * RN + j] = lds[rslot(i, j)];
    __syncthreads();
  } else {
    R* lds = reinterpret_cast<R*>(lds_raw);
#pragma unroll
    for (int i = 0; i < F::E / RP; ++i)
#pragma unroll
      for (int m = 0; m < RP; ++m) lds[wslot(i, m)] = v[i * RP + m].x;
    __syncthreads();
#pragma unroll
    for (int i = 0; i < F::E / RN; ++i)
#pragma unroll
      for (int j = 0; j < RN; ++j) v[i * RN + j].x = lds[rslot(i, j)];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < F::E / RP; ++i)
#pragma unroll
      for (int m = 0; m < RP; ++m) lds[wslot(i, m)] = v[i * RP + m].y;
    __syncthreads();
#pragma unroll
    for (int i = 0; i < F::E / RN; ++i)
#pragma unroll
      for (int j = 0; j < RN; ++j) v[i * RN + j].y = lds[rslot(i, j)];
    __syncthreads();
  }
}

// One Stockham stage on register-resident data.  With FP64 twiddle tables
// (TW = double) and FP32 data, each butterfly is evaluated in FP64 and
// rounded back to FP32 once: single-precision butterflies lose norm
// systematically (about 1e-7 per transform pair), which compounds linearly
// over hundreds of slices, while rounding only the stage outputs grows as a
// random walk.
__device__ __forceinline__ cpx<double> to64(cpx<float> a) { return mk((double)a.x, (double)a.y); }
__device__ __forceinline__ cpx<float> to32(cpx<double> a) { return mk((float)a.x, (float)a.y); }

// Radix-16 butterfly (4 x 4) on FP32 data with each radix-4 layer evaluated in
// FP64 and rounded to FP32 between the layers: FP64 accuracy per operation
// with only four complex doubles live at a time.  T/k: stage twiddles
// W^(j k) applied to input j (TWID = false for the first stage).
template <int DIR, bool TWID>
__device__ __forceinline__ void dft16_l64(cpx<float>* v, const cpx<double>* T, int k) {
  const double c1 = 0.92387953251128675612818318939679;  // cos(pi/8)
  const double s1 = 0.38268343236508977172845998403040;  // sin(pi/8)
  const double h = 0.70710678118654752440084436210485;
  const double d = (double)DIR;
  cpx<float> y[16];
#pragma unroll
  for (int n2 = 0; n2 < 4; ++n2) {
    cpx<double> x[4];
#pragma unroll
    for (int n1 = 0; n1 < 4; ++n1) {
      x[n1] = to64(v[n2 + 4 * n1]);
      if constexpr (TWID) {
        if (n2 + 4 * n1 > 0) x[n1] = twmul<DIR>(x[n1], T[k * 16 + n2 + 4 * n1]);
      }
    }
    dft4<DIR>(x);
    if (n2 == 1) {
      x[1] = cmulcs(x[1], c1, d * s1);
      x[2] = cmulcs(x[2], h, d * h);
      x[3] = cmulcs(x[3], s1, d * c1);
    } else if (n2 == 2) {
      x[1] = cmulcs(x[1], h, d * h);
      x[2] = rot<DIR>(x[2]);
      x[3] = rot<DIR>(cmulcs(x[3], h, d * h));
    } else if (n2 == 3) {
      x[1] = cmulcs(x[1], s1, d * c1);
      x[2] = rot<DIR>(cmulcs(x[2], h, d * h));
      x[3] = cmulcs(x[3], -c1, -d * s1);
    }
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) y[n2 * 4 + k1] = to32(x[k1]);
    // one group's twiddle loads in flight at a time (register pressure)
    if constexpr (TWID) compiler_fence();
  }
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    cpx<double> z[4] = {to64(y[k1]), to64(y[4 + k1]), to64(y[8 + k1]), to64(y[12 + k1])};
    dft4<DIR>(z);
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = to32(z[k2]);
  }
}

template <class F, int S, int DIR, typename R, typename TW>
__device__ __forceinline__ void reg_stage(cpx<R> (&v)[F::E], const StageTw<TW>& tw, int t) {
  constexpr int RR = F::radix(S);
  constexpr int LS = F::stride(S);
  if constexpr (std::is_same<TW, twc>::value) {
    if constexpr (LS > 1) {
      const twc* T = tw.t[S];
#pragma unroll
      for (int i = 0; i < F::E / RR; ++i) {
        const int b = t + i * F::TC;
        const int k = b % LS;
#pragma unroll
        for (int j = 1; j < RR; ++j) v[i * RR + j] = twmul_comp<DIR>(v[i * RR + j], T[k * RR + j]);
      }
    }
#pragma unroll
    for (int i = 0; i < F::E / RR; ++i) dft_comp<RR, DIR>(&v[i * RR]);
  } else if constexpr (sizeof(TW) > sizeof(R) && RR == 16) {
#pragma unroll
    for (int i = 0; i < F::E / 16; ++i) {
      const int b = t + i * F::TC;
      dft16_l64<DIR, (LS > 1)>(&v[i * 16], tw.t[S], b % LS);
    }
  } else if constexpr (sizeof(TW) > sizeof(R)) {
#pragma unroll
    for (int i = 0; i < F::E / RR; ++i) {
      cpx<TW> x[RR];
#pragma unroll
      for (int j = 0; j < RR; ++j) x[j] = mk((TW)v[i * RR + j].x, (TW)v[i * RR + j].y);
      if constexpr (LS > 1) {
        const cpx<TW>* T = tw.t[S];
        const int b = t + i * F::TC;
        const int k = b % LS;
#pragma unroll
        for (int j = 1; j < RR; ++j) x[j] = twmul<DIR>(x[j], T[k * RR + j]);
      }
      dft<RR, DIR>(x);
#pragma unroll
      for (int j = 0; j < RR; ++j) v[i * RR + j] = mk((R)x[j].x, (R)x[j].y);
    }
  } else {
#pragma unroll
    for (int i = 0; i < F::E / RR; ++i) {
      if constexpr (LS > 1) {
        const cpx<TW>* T = tw.t[S];
        const int b = t + i * F::TC;
        const int k = b % LS;
#pragma unroll
        for (int j = 1; j < RR; ++j) v[i * RR + j] = twmul<DIR>(v[i * RR + j], T[k * RR + j]);
      }
      dft<RR, DIR>(&v[i * RR]);
      // one butterfly's twiddles in flight at a time when a thread has several
      if constexpr (F::E > 16 && RR >= 8) compiler_fence();
    }
  }
}

// Full transform on register-resident data.  On entry v holds the stage-0
// inputs (v[i*R0 + j] = x[t + i*TC + j*N/R0]); on exit v holds the outputs
// of the last stage (v[i*RL + m] = X[t + i*TC + m*N/RL]).
template <class F, int DIR, int C, int PS, bool SPLIT, typename R, typename TW>
__device__ __forceinline__ void reg_fft(cpx<R> (&v)[F::E], const StageTw<TW>& tw, void* lds, int t,
                                        int c) {
  reg_stage<F, 0, DIR>(v, tw, t);
  if constexpr (F::NS > 1) {
    reg_exchange<F, 1, C, PS, SPLIT>(v, lds, t, c);
    reg_stage<F, 1, DIR>(v, tw, t);
  }
  if constexpr (F::NS > 2) {
    reg_exchange<F, 2, C, PS, SPLIT>(v, lds, t, c);
    reg_stage<F, 2, DIR>(v, tw, t);
  }
  if constexpr (F::NS > 3) {
    reg_exchange<F, 3, C, PS, SPLIT>(v, lds, t, c);
    reg_stage<F, 3, DIR>(v, tw, t);
  }
}

// Position of register e on entry (stage-0 input layout).
template <class F>
__device__ __forceinline__ int first_pos(int t, int e) {
  constexpr int R0 = F::R0;
  return t + (e / R0) * F::TC + (e % R0) * (F::N / R0);
}
// Position of register e on exit (last-stage output layout).
template <class F>
__device__ __forceinline__ int last_pos(int t, int e) {
  constexpr int RL = F::last_radix();
  return t + (e / RL) * F::TC + (e % RL) * (F::N / RL);
}

// Byte offset of register e relative to the thread's first element, on entry
// (stage-0 layout) and on exit (last-stage layout); es_b = element stride in bytes.
template <class F>
__device__ __forceinline__ uint32_t first_off(int e, uint32_t es_b) {
  return (uint32_t)((e / F::R0) * F::TC + (e % F::R0) * (F::N / F::R0)) * es_b;
}
template <class F>
__device__ __forceinline__ uint32_t last_off(int e, uint32_t es_b) {
  constexpr int RL = F::last_radix();
  return (uint32_t)((e / RL) * F::TC + (e % RL) * (F::N / RL)) * es_b;
}

// Plain single-direction transform along one axis, C sequences per WG.
// Batch b starts at in + b*bs (elements); sequence q of it at q*ss_b bytes,
// its element p at p*es_b bytes; nbytes bounds one batch.
// TW: twiddle / butterfly arithmetic (R, or twc: FP32 data with hi+lo
// compensated constants and twiddles).
template <typename R, class F, int DIR, int C, int PS, bool SPLIT, typename TW = R>
__global__ void __launch_bounds__(F::TC * C) fft_reg_kernel(const cpx<R>* in, cpx<R>* out,
                                                            uint32_t es_b, uint32_t ss_b,
                                                            long long bs, uint32_t nbytes,
                                                            int nseq, StageTw<TW> tw, R scale) {
  __shared__ unsigned char lds[LdsSize<F, C, PS>::slots * (SPLIT ? sizeof(R) : sizeof(cpx<R>))];
  const int tid = threadIdx.x;
  const int c = tid % C;
  const int t = tid / C;
  const int q = xcd_tile(blockIdx.x, gridDim.x) * C + c;
  const bool live = q < nseq;
  const Gio<R> gi(in + (long long)blockIdx.y * bs, nbytes);
  const Gio<R> go(out + (long long)blockIdx.y * bs, nbytes);
  const uint32_t vo = (uint32_t)(live ? q : 0) * ss_b + (uint32_t)t * es_b;
  cpx<R> v[F::E];
#pragma unroll
  for (int e = 0; e < F::E; ++e) v[e] = live ? gi.ld(vo, first_off<F>(e, es_b)) : mk(R(0), R(0));
  reg_fft<F, DIR, C, PS, SPLIT>(v, tw, lds, t, c);
  if (live) {
#pragma unroll
    for (int e = 0; e < F::E; ++e) go.st(cscale(v[e], scale), vo, last_off<F>(e, es_b));
  }
}

}  // namespace rsim
