// Instantiation of the two-role LU pair kernel for (double source, float factors).
#include "bcr_factor_tw.inc"

namespace bcr {
template bool launch_pair_lu_tw<double, float>(FactorArgs<double, float> const&, hipStream_t, hipError_t*);
template bool launch_pair_lu_tw_batched<double, float>(BatchTable<FactorArgs<double, float>> const&, int, int, int, int, hipStream_t,
                                              hipError_t*);
}  // namespace bcr
