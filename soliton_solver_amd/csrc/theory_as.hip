// Kernel instantiations for the anisotropic s+id superconductor.
#include "relax_kernels.h"
#include "theory_as.h"

namespace ss {
SS_DEFINE_THEORY_OPS(AnisotropicSC, ops_anisotropic_sc)
}  // namespace ss
