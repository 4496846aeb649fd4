// Kernel instantiations for the Ginzburg-Landau superconductor (Abelian Higgs).
#include "relax_kernels.h"
#include "theory_gl.h"

namespace ss {
SS_DEFINE_THEORY_OPS(GinzburgLandau, ops_ginzburg_landau)
}  // namespace ss
