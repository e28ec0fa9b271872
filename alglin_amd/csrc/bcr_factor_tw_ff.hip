// Instantiation of the two-role LU pair kernel for (float source, float factors).
#include "bcr_factor_tw.inc"

namespace bcr {
template bool launch_pair_lu_tw<float, float>(FactorArgs<float, float> const&, hipStream_t, hipError_t*);
template bool launch_pair_lu_tw_batched<float, float>(BatchTable<FactorArgs<float, float>> const&, int, int, int, int, hipStream_t,
                                              hipError_t*);
}  // namespace bcr
