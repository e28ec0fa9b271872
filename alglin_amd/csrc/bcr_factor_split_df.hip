// Instantiation of the two-wave LU pair kernel for (double source, float factors).
#include "bcr_factor_split.inc"

namespace bcr {
template bool launch_pair_lu_split<double, float>(FactorArgs<double, float> const&, hipStream_t, hipError_t*);
}  // namespace bcr
