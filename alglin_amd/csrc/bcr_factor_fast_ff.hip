// Instantiation of the register-resident LU pair kernel for (float source, float factors).
#include "bcr_factor_fast.inc"

namespace bcr {
template bool launch_pair_lu_fast<float, float>(FactorArgs<float, float> const&, hipStream_t, hipError_t*);
}  // namespace bcr
