// Kernel instantiations for the chiral magnet.
#include "relax_kernels.h"
#include "theory_chiral.h"

namespace ss {
SS_DEFINE_THEORY_OPS(ChiralMagnet, ops_chiral_magnet)
}  // namespace ss
