// Kernel instantiations for the ferromagnetic superconductor.
#include "relax_kernels.h"
#include "theory_fs.h"

namespace ss {
SS_DEFINE_THEORY_OPS(FerromagneticSC, ops_fs)
}  // namespace ss
