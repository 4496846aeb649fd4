// Kernel instantiations for the spin-triplet superconducting ferromagnet.
#include "relax_kernels.h"
#include "theory_st.h"

namespace ss {
SS_DEFINE_THEORY_OPS(SpinTripletSFM, ops_st)
}  // namespace ss
