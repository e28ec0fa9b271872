// Host-callable launchers of the device kernels (defined in the .hip files).
#pragma once

#include "bcr_common.hh"

namespace bcr {

template <typename TW, typename TX> struct SolveArgs;
template <typename TW, typename TA> struct AssembleArgs;
template <typename TA> struct MvArgs;

// Householder QR / QRP pair elimination for every n <= 64, nr <= 16: T in
// LDS, the right-hand columns in chunks of cw next to it (pair_qr_plan picks
// cw and the LDS bytes); false for LU or a shape beyond
bool pair_qr_plan(int n, int nr, int nx, size_t elem, size_t lds_limit, int* cw, size_t* bytes);
template <typename TS, typename TW>
bool launch_pair_qr(int method, FactorArgs<TS, TW> const& a, size_t lds_limit, hipStream_t st, hipError_t* err);
template <typename TS, typename TW>
hipError_t launch_pair_generic(int method, FactorArgs<TS, TW> const& a, TW* gscratch, size_t lds_limit,
                               hipStream_t st);
template <typename TS, typename TW>
bool launch_pair_lu_fast(FactorArgs<TS, TW> const& a, hipStream_t st, hipError_t* err);
template <typename TS, typename TW>
bool launch_pair_lu_split(FactorArgs<TS, TW> const& a, hipStream_t st, hipError_t* err);
// register-tiled LU for any n <= 64 (2n + nr <= 160, 3n + nx <= 208)
bool pair_lu_tiled_fits(int n, int nr, int nx);
template <typename TS, typename TW>
bool launch_pair_lu_tiled(FactorArgs<TS, TW> const& a, hipStream_t st, hipError_t* err);
// column-split waves LU for n <= 64 (n <= 48 with nx <= 4 and 2n + nr <= 128,
// else nx <= 16 and 2n + nr <= 192); false when the shape is not covered
bool pair_lu_cols_fits(int n, int nr, int nx);
template <typename TS, typename TW>
bool launch_pair_lu_cols(FactorArgs<TS, TW> const& a, hipStream_t st, hipError_t* err);
// the matrix-core pair: one wave factors T, the right-hand columns are
// blocked triangular solves and GEMMs on the matrix cores.  FP32 factors:
// n <= 64, n + nr <= 80, nr <= 8, nx <= 16; FP64 factors: n <= 48,
// n + nr <= 64; false otherwise
bool pair_lu_mfma_fits(int n, int nr, int nx, bool fp64);
template <typename TS, typename TW>
bool launch_pair_lu_mfma(FactorArgs<TS, TW> const& a, hipStream_t st, hipError_t* err);

template <typename TW, typename TX>
hipError_t launch_forward(int method, SolveArgs<TW, TX> const& a, hipStream_t st);
template <typename TW, typename TX>
hipError_t launch_backward(int method, SolveArgs<TW, TX> const& a, hipStream_t st);
// streamed LU sweeps, one wavefront per pair (n <= 64)
template <typename TW, typename TX>
bool launch_sweep_lu_stream(bool forward, SolveArgs<TW, TX> const& a, hipStream_t st, hipError_t* err);
// wave-synchronous LU sweeps (n <= 32); false when the shape is not covered
template <typename TW, typename TX>
bool launch_sweep_lu_fast(bool forward, SolveArgs<TW, TX> const& a, hipStream_t st, hipError_t* err);
template <typename TP, typename TD>
hipError_t launch_reduce(const TP* part, int first, int count, int m, TP* tmp, int max_chunks, TD* dst, int mrow,
                         int64_t ldd, hipStream_t st);

template <typename TW, typename TA>
hipError_t launch_assemble_last(AssembleArgs<TW, TA> const& a, hipStream_t st);
hipError_t launch_lu_last(double* Hm, int Nc, int32_t* ipiv, int32_t* status, size_t lds_limit, hipStream_t st);
template <typename TX>
hipError_t launch_last_solve(LastSolveArgs<TX> const& a, size_t lds_limit, hipStream_t st);
hipError_t launch_lupq_last(double* Hm, int Nc, int32_t* P, int32_t* Q, int32_t* status, size_t lds_limit,
                            hipStream_t st);
hipError_t launch_svd_last(SvdArgs const& a, hipStream_t st);
hipError_t launch_qr_last(bool pivot, double* Hm, int Nr, int Nc, double* tau, int32_t* jpvt, int32_t* status,
                          size_t lds_limit, hipStream_t st);

// r = b - A x, y = A x, or (absval) |A| |x| + |b|
template <typename TA>
hipError_t launch_mv(MvArgs<TA> const& a, double* Csum, double* tmp, int max_chunks, hipStream_t st,
                     bool absval = false);
// *out = max_i |r_i| / den_i (device scalar)
hipError_t launch_max_ratio(const double* r, const double* den, int64_t len, double* tmp, int max_chunks, double* out,
                            hipStream_t st);
hipError_t launch_nrm2(const double* v, int64_t len, double* tmp, int max_chunks, double* out, hipStream_t st);

// dst[i] = TD(src[i]) for i < count  (also a zero-fill when src == nullptr)
template <typename TS, typename TD>
hipError_t launch_convert(const TS* src, TD* dst, int64_t count, hipStream_t st);
// dst[i] += TD(alpha * src[i])
hipError_t launch_axpy(double alpha, const double* src, double* dst, int64_t count, hipStream_t st);

}  // namespace bcr
