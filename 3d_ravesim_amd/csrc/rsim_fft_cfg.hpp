// Compile-time FFT schedules for the register-resident kernels.
//
// Each hot size has one radix sequence (shared by row and column kernels, so
// the per-stage twiddle tables are shared too) with first radix == last radix:
// an inverse transform's outputs then sit in the registers where the
// following forward transform expects its inputs, which lets the fused
// multislice kernels chain IFFT -> pointwise -> FFT without an LDS exchange.
#pragma once

#include "rsim_fft.hpp"

namespace rsim {

template <int N>
struct Sched;
template <> struct Sched<256> { using F = RegFac<256, 16, 16, 16>; };
template <> struct Sched<512> { using F = RegFac<512, 16, 8, 8, 8>; };
template <> struct Sched<1024> { using F = RegFac<1024, 16, 16, 4, 16>; };
template <> struct Sched<2048> { using F = RegFac<2048, 16, 16, 8, 16>; };
template <> struct Sched<4096> { using F = RegFac<4096, 16, 16, 16, 16>; };
template <> struct Sched<8192> { using F = RegFac<8192, 16, 16, 2, 16, 16>; };
// 2^a * 3 lengths: E = 24 elements per thread (a multiple of 8, 6 and 12)
template <> struct Sched<3072> { using F = RegFac<3072, 24, 8, 6, 8, 8>; };
template <> struct Sched<6144> { using F = RegFac<6144, 24, 8, 12, 8, 8>; };
// row pass of 16384-point axes: 32 elements per thread (512 threads per row),
// three stages (two LDS exchanges per transform)
template <> struct Sched<16384> { using F = RegFac<16384, 32, 32, 16, 32>; };
// sub-transform of the three-pass 16384-point column schedule (128 x 128)
template <> struct Sched<128> { using F = RegFac<128, 8, 8, 2, 8>; };

// Launch shape along the contiguous axis (rows) and the strided axis
// (columns).  Columns put C neighbouring columns in one workgroup so every
// wavefront load covers C*sizeof(complex) contiguous bytes per row.
template <typename R, int N>
struct RowShape {
  using F = typename Sched<N>::F;
  static constexpr int C = (256 / F::TC) > 0 ? (256 / F::TC) : 1;
  static constexpr bool SPLIT = (size_t)F::N * C * sizeof(cpx<R>) > 65536;
  static constexpr int PS = pad_code(4 + (C >= 2) + (C >= 4) + (C >= 8) + (C >= 16), C);
};
template <typename R, int N>
struct ColShape {
  using F = typename Sched<N>::F;
  static constexpr int Cmax = 1024 / F::TC;
  static constexpr int Cw = sizeof(R) == 4 ? 16 : 8;   // columns wanted
  static constexpr int Clds0 = 65536 / (N * (int)sizeof(R));  // split exchange <= 64 KiB
  static constexpr int Clds = Clds0 > 0 ? Clds0 : 1;
  static constexpr int C1 = Cmax < Cw ? Cmax : Cw;
  static constexpr int C = C1 < Clds ? C1 : Clds;
  static constexpr bool SPLIT = true;
  static constexpr int PS = pad_code(4 + (C >= 2) + (C >= 4) + (C >= 8) + (C >= 16), C);
};

// Host-side description of a schedule (for building twiddle tables).
struct SchedInfo {
  int N;
  int nst;
  int radix[4];
  int E;
};

inline bool sched_info(int N, SchedInfo* si) {
#define RSIM_SI(n)                                               \
  case n: {                                                      \
    using F = Sched<n>::F;                                       \
    si->N = n;                                                   \
    si->nst = F::NS;                                             \
    si->radix[0] = F::R0;                                        \
    si->radix[1] = F::R1;                                        \
    si->radix[2] = F::R2;                                        \
    si->radix[3] = F::R3;                                        \
    si->E = F::E;                                                \
    return true;                                                 \
  }
  switch (N) {
    RSIM_SI(256)
    RSIM_SI(512)
    RSIM_SI(1024)
    RSIM_SI(2048)
    RSIM_SI(4096)
    RSIM_SI(8192)
    RSIM_SI(3072)
    RSIM_SI(6144)
    RSIM_SI(16384)
    RSIM_SI(128)
    default: return false;
  }
#undef RSIM_SI
}

}  // namespace rsim
