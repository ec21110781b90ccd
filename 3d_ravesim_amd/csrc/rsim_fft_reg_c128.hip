// Plain register-resident FFT kernels, complex128.
#include "rsim_fft_reg_impl.hpp"

namespace rsim {
bool launch_reg_fft_c128(int dir, int N, bool colmode, const void* in, void* out, long long es,
                         long long ss, long long bs, int nseq, int batch, const void* const* tw,
                         double scale, hipStream_t s) {
  return launch_reg<double>(dir, N, colmode, in, out, es, ss, bs, nseq, batch, tw, scale, s);
}
}  // namespace rsim
