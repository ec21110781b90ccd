// Fused multislice row kernels for length-16384 axes whose second transform
// recomputes its addresses (the LD launch-shape variants), kept out of the
// base variants' translation unit.  Columns of that length use the
// three-pass schedule.
#include "rsim_fused.hpp"

namespace rsim {
RSIM_FUSED_INSTANTIATE_ROWS_LD(16384)
}  // namespace rsim
