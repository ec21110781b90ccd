// Fused multislice kernels for length-3072 axes whose second transform
// recomputes its addresses (the LD launch-shape variants), kept out of the
// base variants' translation unit.
#include "rsim_fused.hpp"

namespace rsim {
RSIM_FUSED_INSTANTIATE_LD(3072)
}  // namespace rsim
