// Kernel instantiations for the chiral liquid crystal.
#include "relax_kernels.h"
#include "theory_lc.h"

namespace ss {
SS_DEFINE_THEORY_OPS(LiquidCrystal, ops_lc)
}  // namespace ss
