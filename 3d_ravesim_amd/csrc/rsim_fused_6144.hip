// Fused multislice kernels for length-6144 axes (2^a * 3: radix-8 and prime-factor
// radix-12 stages, 24 elements per thread).
#include "rsim_fused.hpp"

namespace rsim {
RSIM_FUSED_INSTANTIATE(6144)
}  // namespace rsim
