// Instantiation of the register-resident LU pair kernel for (double source, double factors).
#include "bcr_factor_fast.inc"

namespace bcr {
template bool launch_pair_lu_fast<double, double>(FactorArgs<double, double> const&, hipStream_t, hipError_t*);
template bool launch_pair_lu_fast_batched<double, double>(BatchTable<FactorArgs<double, double>> const&, int, int, int, int, hipStream_t,
                                              hipError_t*);
}  // namespace bcr
