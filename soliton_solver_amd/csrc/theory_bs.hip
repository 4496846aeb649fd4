// Kernel instantiations for the baby Skyrme model.
#include "relax_kernels.h"
#include "theory_bs.h"

namespace ss {
SS_DEFINE_THEORY_OPS(BabySkyrme, ops_baby_skyrme)
}  // namespace ss
