// Fused multislice kernels for length-6144 axes whose second transform
// recomputes its addresses (the LD launch-shape variants), kept out of the
// base variants' translation unit.
#include "rsim_fused.hpp"

namespace rsim {
RSIM_FUSED_INSTANTIATE_LD(6144)
}  // namespace rsim
