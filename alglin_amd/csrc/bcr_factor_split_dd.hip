// Instantiation of the two-wave LU pair kernel for (double source, double factors).
#include "bcr_factor_split.inc"

namespace bcr {
template bool launch_pair_lu_split<double, double>(FactorArgs<double, double> const&, hipStream_t, hipError_t*);
template bool launch_pair_lu_split_batched<double, double>(BatchTable<FactorArgs<double, double>> const&, int, int, int, int, hipStream_t,
                                              hipError_t*);
}  // namespace bcr
