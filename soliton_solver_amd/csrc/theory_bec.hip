// Kernel instantiations for the rotating Bose-Einstein condensate.
#include "relax_kernels.h"
#include "theory_bec.h"

namespace ss {
SS_DEFINE_THEORY_OPS(BoseEinsteinCondensate, ops_bec)
}  // namespace ss
