// Instantiation of the two-role LU pair kernel for (double source, double factors).
#include "bcr_factor_tw.inc"

namespace bcr {
template bool launch_pair_lu_tw<double, double>(FactorArgs<double, double> const&, hipStream_t, hipError_t*);
template bool launch_pair_lu_tw_batched<double, double>(BatchTable<FactorArgs<double, double>> const&, int, int, int, int, hipStream_t,
                                              hipError_t*);
}  // namespace bcr
