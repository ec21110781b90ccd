// Fused multislice kernels for length-4096 axes (one translation unit per size
// so the kernel variants compile in parallel).
#include "rsim_fused.hpp"

namespace rsim {
RSIM_FUSED_INSTANTIATE(4096)
}  // namespace rsim
