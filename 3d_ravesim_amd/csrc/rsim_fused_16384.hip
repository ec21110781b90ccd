// Fused multislice kernels for length-16384 axes: the register row pass
// (1024 threads per row) and the three-pass column schedule (16384 = 128 x 128).
#include "rsim_fused.hpp"

namespace rsim {

RSIM_FUSED_INSTANTIATE_ROWS(16384)

namespace {
// C neighbouring columns per tile: every row segment is C * 8 contiguous
// bytes (C = 16: one 128-byte line).  Wider tiles read longer runs of each
// DRAM page per workgroup.
template <int C_>
struct BigShape {
  using F = Sched<128>::F;
  static constexpr int C = C_;
  static constexpr bool SPLIT = false;   // 128 x C x 8 B <= 64 KiB
  static constexpr int PS = pad_code(4 + (C >= 2) + (C >= 4) + (C >= 8) + (C >= 16) + (C >= 32) +
                                         (C >= 64),
                                     C);
};

template <typename TW, int C>
bool big_col_go_impl(void* buf, int nx, int ny, const StageTw<TW>& tw, const BigTwElem<TW>* wN,
                     const double2* Bx, const double2* Ay, const int* Rmax, hipStream_t s) {
  using S = BigShape<C>;
  using F = typename S::F;
  constexpr int N1 = F::N, N2 = F::N;
  if (ny != N1 * N2) return false;
  const unsigned long long bytes = (unsigned long long)nx * ny * 8ull;
  if (bytes > 0xFFFFFFFFull) return false;
  const uint32_t nb = (uint32_t)bytes;
  cpx<float>* b = reinterpret_cast<cpx<float>*>(buf);
  const dim3 block(F::TC * S::C);
  const int tiles = (nx + S::C - 1) / S::C;
  hipLaunchKernelGGL((ms_colAC_kernel<F, S::C, S::PS, S::SPLIT, -1, TW>), dim3(tiles, N2), block,
                     0, s, b, nx, N2, nb, tw, wN);
  hipLaunchKernelGGL((ms_colB_kernel<F, S::C, S::PS, S::SPLIT, TW>), dim3(tiles, N1), block, 0, s,
                     b, nx, N1, nb, tw, wN, Bx, Ay, Rmax);
  hipLaunchKernelGGL((ms_colAC_kernel<F, S::C, S::PS, S::SPLIT, 1, TW>), dim3(tiles, N2), block,
                     0, s, b, nx, N2, nb, tw, wN);
  return true;
}

template <typename TW>
bool big_col_dispatch(int v, void* buf, int nx, int ny, const StageTw<TW>& tw,
                      const BigTwElem<TW>* wN, const double2* Bx, const double2* Ay,
                      const int* Rmax, hipStream_t s) {
  switch (v) {
    case 0: return big_col_go_impl<TW, 16>(buf, nx, ny, tw, wN, Bx, Ay, Rmax, s);
    case 1: return big_col_go_impl<TW, 32>(buf, nx, ny, tw, wN, Bx, Ay, Rmax, s);
    case 2: return big_col_go_impl<TW, 64>(buf, nx, ny, tw, wN, Bx, Ay, Rmax, s);
    default: return false;
  }
}
}  // namespace

template <>
bool big_col_go<float>(int v, void* buf, int nx, int ny, const StageTw<float>& tw,
                       const BigTwElem<float>* wN, const double2* Bx, const double2* Ay,
                       const int* Rmax, hipStream_t s) {
  return big_col_dispatch<float>(v, buf, nx, ny, tw, wN, Bx, Ay, Rmax, s);
}
template <>
bool big_col_go<double>(int v, void* buf, int nx, int ny, const StageTw<double>& tw,
                        const BigTwElem<double>* wN, const double2* Bx, const double2* Ay,
                        const int* Rmax, hipStream_t s) {
  return big_col_dispatch<double>(v, buf, nx, ny, tw, wN, Bx, Ay, Rmax, s);
}
template <>
bool big_col_go<twc>(int v, void* buf, int nx, int ny, const StageTw<twc>& tw,
                     const BigTwElem<twc>* wN, const double2* Bx, const double2* Ay,
                     const int* Rmax, hipStream_t s) {
  return big_col_dispatch<twc>(v, buf, nx, ny, tw, wN, Bx, Ay, Rmax, s);
}

}  // namespace rsim
