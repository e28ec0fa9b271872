// Instantiation of the register-resident LU pair kernel for (double source, float factors).
#include "bcr_factor_fast.inc"

namespace bcr {
template bool launch_pair_lu_fast<double, float>(FactorArgs<double, float> const&, hipStream_t, hipError_t*);
template bool launch_pair_lu_fast_batched<double, float>(BatchTable<FactorArgs<double, float>> const&, int, int, int, int, hipStream_t,
                                              hipError_t*);
}  // namespace bcr
