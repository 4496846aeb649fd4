// Kernel instantiations for the Maxwell-Chern-Simons-Higgs (anyon) model.
#include "relax_kernels.h"
#include "theory_mcsh.h"

namespace ss {
SS_DEFINE_THEORY_OPS(MaxwellChernSimonsHiggs, ops_mcsh)
}  // namespace ss
