// Argument records and launchers of the batched (many systems per launch)
// kernels.  Each record's pointers are already offset to one system's slice
// of the packed arrays.
#pragma once

#include "bcr_launch.hh"

namespace bcr {

template <typename TA, typename TW>
struct PrepArgs {
  int N, n, nr, nx;
  const TA *D, *E, *B, *C;
  TW *wD, *wE, *wB, *wC;
};

// dst[(i % mrow) + (i / mrow) * ldd] += sum_{k=first}^{first+count-1} part[k*m + i]
template <typename TP, typename TD>
struct ReduceArgs {
  const TP* part;
  int first, count, m, mrow;
  int64_t ldd;
  TD* dst;
  const int32_t* skip = nullptr;   // batched: the system is done when *skip != 0
};

struct NormArgs {
  const double* b;
  const double* r;
  int64_t len;
  int sys;   // output slot: out[2*sys], out[2*sys+1]
  const int32_t* skip = nullptr;   // the system is done when *skip != 0
};

// synthetic system s: D_k = shift I + G/sqrt(n), E_k = G/sqrt(n), B, C, H, Cq, F = G;
// with sing_k >= 1 the pair column block [E_{k-1}; D_k] (k = sing_k) is made
// rank deficient up to eps: its last column becomes the mean of its first two
// plus eps G/sqrt(n)
struct GenArgs {
  int sys, N, n, qr, qx, nr, nx, sing_k;
  double shift, eps;
  double *D, *E, *B, *C, *H, *Cq, *F;
};

// per-system features of the LU-vs-QRP selector (12 floats at out[12*sys]),
// from the factored pair panels, the factored last block and the residual
// of the first solve.  Two launches: panel_stats (one wave per pair panel,
// every system at once) writes (min |u_ii|, max |u_ii|, max_{i<=k} |u_ik|)
// of each panel to pstat; features (one workgroup per system) reduces its
// panels' triples and adds the last-block and residual terms.
template <typename TW>
struct FeatArgs {
  int N, n, qr, nr, Nc, sys;
  const TW* Tm;          // N panels of tsize(n)
  const double* Hm;      // factored Nc x Nc last block
  const double* norms;   // norms[2*sys] = ||b||, norms[2*sys+1] = ||r||
  float* pstat;          // 3 floats per panel j = 1 .. N-1 of this system
};

struct CopyArgs {
  const double* src;
  double* dst;
  int64_t len;
  const int32_t* skip = nullptr;   // the system is done when *skip != 0
};

// panels: BatchTable over workgroups of four panels (workgroup b of record e
// covers panels j = 1 + 4 (b - first[e]) + w, w < 4)
template <typename TW>
hipError_t launch_panel_stats_batched(BatchTable<FeatArgs<TW>> const& t, int nblocks, hipStream_t st);
template <typename TW>
hipError_t launch_features_batched(const FeatArgs<TW>* args, int nsys, float* out, hipStream_t st);
hipError_t launch_copy_batched(const CopyArgs* args, int nsys, hipStream_t st);
// dst_s += src_s per record (the same CopyArgs records)
hipError_t launch_add_batched(const CopyArgs* args, int nsys, hipStream_t st);

// One refinement test per system, on the device: rel = ||r|| / ||b|| from
// norms[2s], norms[2s+1]; a system not yet done becomes done when it has
// converged (rel <= tol), is not finite, its factorization failed
// (status != 0) or it stalled (three sweeps in a row without halving its best
// residual); with giveup > 0 also when, at its last contraction rate, it
// would need more than giveup further sweeps (a better solver takes over);
// *open counts the systems left open (cleared by the caller).
struct ConvergeArgs {
  const double* norms;
  const int32_t* status;
  double* best;
  int32_t* stagnant;
  int32_t* done;
  double* rel;
  int32_t* sweeps;
  int32_t* open;
  double tol;
  int it;      // sweep index; 0 resets best / stagnant
  int nsys;
  int giveup;  // 0: off
};
hipError_t launch_converge_batched(ConvergeArgs const& a, hipStream_t st);

template <typename TA, typename TW>
hipError_t launch_prep_batched(const PrepArgs<TA, TW>* args, int nsys, hipStream_t st);
template <typename TP, typename TD>
hipError_t launch_reduce_batched(const ReduceArgs<TP, TD>* args, int nsys, hipStream_t st);
hipError_t launch_norms_batched(const NormArgs* args, int nsys, double* out, hipStream_t st);
// out[sys] = max_i |r_i| / b_i over the system's rows (b = |A||x| + |b|:
// the componentwise backward error; 0 / 0 counts as 0)
hipError_t launch_berr_batched(const NormArgs* args, int nsys, double* out, hipStream_t st);
hipError_t launch_gen_batched(const GenArgs* args, int nsys, uint64_t seed, hipStream_t st);

// pair elimination (factor), one launch per reduction level and kernel class
inline int pair_lu_fast_ppb(int n, int nx, int nr) {
  if (n <= 4 && nx <= 4 && nr <= 4) return kWave / 8;
  if (n <= 8 && nx <= 8 && nr <= 8) return kWave / 16;
  if (n <= 16 && nx <= 8 && nr <= 8) return kWave / 32;
  if (n <= 32 && nx <= 8 && nr <= 8) return 1;
  return 0;
}
template <typename TS, typename TW>
bool launch_pair_lu_fast_batched(BatchTable<FactorArgs<TS, TW>> const& t, int nblocks, int n, int nx, int nr,
                                 hipStream_t st, hipError_t* err);
template <typename TS, typename TW>
bool launch_pair_lu_split_batched(BatchTable<FactorArgs<TS, TW>> const& t, int nblocks, int n, int nx, int nr,
                                  hipStream_t st, hipError_t* err);
template <typename TS, typename TW>
bool launch_pair_lu_tiled_batched(BatchTable<FactorArgs<TS, TW>> const& t, int nblocks, int n, int nx, int nr,
                                  hipStream_t st, hipError_t* err);
// waves per pair: 0 default, else 4 / 8 / 16 (n <= 48) or 8 / 16 (n <= 64)
template <typename TS, typename TW>
bool launch_pair_lu_cols_batched(BatchTable<FactorArgs<TS, TW>> const& t, int nblocks, int n, int nx, int nr,
                                 hipStream_t st, hipError_t* err, int waves = 0);
template <typename TS, typename TW>
bool launch_pair_lu_mfma_batched(BatchTable<FactorArgs<TS, TW>> const& t, int nblocks, int n, int nx, int nr,
                                 hipStream_t st, hipError_t* err);
template <typename TS, typename TW>
bool launch_pair_qr_batched(int method, BatchTable<FactorArgs<TS, TW>> const& t, int nblocks, int n, int nr, int nx,
                            size_t lds_limit, hipStream_t st, hipError_t* err);
template <typename TS, typename TW>
hipError_t launch_pair_generic_batched(int method, BatchTable<FactorArgs<TS, TW>> const& t, int nblocks, size_t bytes,
                                       size_t lds_limit, hipStream_t st);

// sweeps
int sweep_lu_fast_pairs_per_block(int n);
template <typename TW, typename TX>
bool launch_sweep_lu_fast_batched(bool forward, BatchTable<SolveArgs<TW, TX>> const& t, int nblocks, int n, int nrhs,
                                  hipStream_t st, hipError_t* err);
template <typename TW, typename TX>
bool launch_sweep_lu_stream_batched(bool forward, BatchTable<SolveArgs<TW, TX>> const& t, int nblocks, int n,
                                    hipStream_t st, hipError_t* err);
template <typename TW, typename TX>
hipError_t launch_sweep_generic_batched(bool forward, int method, BatchTable<SolveArgs<TW, TX>> const& t, int nblocks,
                                        int n, int nx, hipStream_t st);

// last block, one workgroup per system
template <typename TW, typename TA>
hipError_t launch_assemble_last_batched(const AssembleArgs<TW, TA>* args, int nsys, hipStream_t st);
hipError_t launch_lu_last_batched(const LastSolveArgs<double>* args, int32_t* const* status, int nsys, int Ncmax,
                                  size_t lds_limit, hipStream_t st);
template <typename TX>
hipError_t launch_last_solve_batched(const LastSolveArgs<TX>* args, int nsys, int Nrmax, int Ncmax, size_t lds_limit,
                                     hipStream_t st);
// one wavefront per system: square last blocks, LU, one right-hand side
// (Ncmax <= 256; false when it does not apply)
template <typename TX>
bool launch_last_solve_lu_wave_batched(const LastSolveArgs<TX>* args, int nsys, int Ncmax, hipStream_t st,
                                       hipError_t* err);

// residual SpMV
template <typename TA>
hipError_t launch_mv_rows_batched(BatchTable<MvArgs<TA>> const& t, int nblocks, hipStream_t st, bool absval = false);
template <typename TA>
hipError_t launch_mv_tail_batched(const MvArgs<TA>* args, const double* const* Csum, int nsys, hipStream_t st,
                                  bool absval = false);

}  // namespace bcr
