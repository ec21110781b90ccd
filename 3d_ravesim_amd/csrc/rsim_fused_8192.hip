// Fused multislice kernels for length-8192 axes (one translation unit per size
// so the kernel variants compile in parallel).
#include "rsim_fused.hpp"

namespace rsim {
RSIM_FUSED_INSTANTIATE(8192)
}  // namespace rsim
