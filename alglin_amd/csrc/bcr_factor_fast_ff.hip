// Instantiation of the register-resident LU pair kernel for (float source, float factors).
#include "bcr_factor_fast.inc"

namespace bcr {
template bool launch_pair_lu_fast<float, float>(FactorArgs<float, float> const&, hipStream_t, hipError_t*);
template bool launch_pair_lu_fast_batched<float, float>(BatchTable<FactorArgs<float, float>> const&, int, int, int, int, hipStream_t,
                                              hipError_t*);
}  // namespace bcr
