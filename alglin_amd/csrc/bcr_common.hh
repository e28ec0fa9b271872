// Shared device-side definitions for the bordered cyclic-reduction kernels.
//
// Data layout in HBM (one system): the block arrays keep the block-contiguous,
// column-major layout of the reference (BABD_BorderedCR.hxx:972-998):
//   D_k, E_k : n x n      at D + k*n*n        (k = 0..N-1)
//   B_k      : n x nx     at B + k*n*nx
//   C_k      : nr x n     at C + k*nr*n       (k = 0..N)
//   T_j      : 2n x n panel (L\U or Householder vectors) + n taus at Tm + j*(2n*n+n)
//   Perm_j   : n ints (LU: 1-based ipiv; QRP: 0-based exchange list)
// so a workgroup that owns a pair streams each block as one contiguous run.
//
// Cross-pair writes of the serial reference (pair j updates C_{j+k}, which
// pair j+2k also updates as its C_jp, BABD_BorderedCR.cc:1124-1143; every
// pair accumulates into F, cc:1146-1154) are made race-free with
//   * two pending-C slots per block, Cpend[slot][x], written at level k by the
//     pair whose jpp == x into slot log2(k)&1 and folded in by the pair that
//     next reads C_x (see bcr_factor.hip for the proof of disjointness), and
//   * one F partial per eliminated block, Fpart[j], summed after the levels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bcr {

constexpr int kWave = 64;
constexpr int kRhsChunk = 16;  // right-hand sides per pass of the sweep kernels

enum Method : int { kLU = 0, kQR = 1, kQRP = 2 };

// Per-level arguments of the pair-elimination kernels.
template <typename TS, typename TW>
struct FactorArgs {
  int N, n, nr, nx;
  int L;           // level stride k (1, 2, 4, ...)
  int npairs;      // pairs at this level
  int slot_cur;    // pending-C slot written at this level
  int first;       // 1 at the first level (nothing pending to fold in yet)
  const TS* sD; const TS* sE; const TS* sB; const TS* sC;  // sources
  TW* D; TW* E; TW* B; TW* C;                              // working blocks
  TW* Tm; int32_t* perm;
  TW* Cpend;       // [2][N+1][nr*n]
  TW* Fpart;       // [N][nr*nx]
  int32_t* status; // set to 1 when a pivot/diagonal is exactly zero
  // partitioned elimination order: (jp, j, jpp) of pair p at pairs[3p..3p+2];
  // nullptr = pure cyclic reduction over [0, N) at stride L
  const int32_t* pairs = nullptr;
  // FP32 LU pairs of a batch: the selector's statistics of panel j (min and
  // max |U_cc|, max |U_rc| over r <= c of its top n x n block) at
  // pstat[3 (j - 1) ..], written by the pair kernels that can; nullptr = none
  float* pstat = nullptr;
};

// Per-level arguments of the forward / backward sweeps.
template <typename TW, typename TX>
struct SolveArgs {
  int N, n, nr, nx, qr, qx;
  int L, npairs;
  int nrhs;
  int64_t ldx;        // leading dimension of X
  TX* x;              // column-major X (ldx x nrhs)
  const TW* D; const TW* E; const TW* B; const TW* C;
  const TW* Tm; const int32_t* perm;
  TX* Xpart;          // [N][nr*nrhs]
  const int32_t* skip = nullptr;   // batched: the system is done when *skip != 0
  const int32_t* pairs = nullptr;  // as FactorArgs::pairs
  // back substitution of QRP pairs (streamed LU sweep): after U^-1, the
  // pair's column exchanges (perm: 0-based exchange list) act on x_j
  int exch = 0;
};

// Arguments of the last-block assembly.
template <typename TW, typename TA>
struct AssembleArgs {
  int N, n, qr, qx, nr, nx, Nr, Nc;
  const TW* D0; const TW* E0; const TW* B0;   // working row 0
  const TW* C;                                // working C (C_0 and C_N used)
  TW const* Cpend;                            // [2][N+1][nr*n]
  const TA* H; const TA* Cq;                  // pristine
  const double* Feff;                         // nr x nx (pristine F + sum of partials)
  double* Hm;                                 // Nr x Nc output
};

// Last-block factorization methods on the GPU (the select_last choices
// LU, LUPQ, QR, QRP of BABD_BorderedCR.hxx:118-127, same numbering).
// kLastSVD stands for SVD, LSS, LSY and PINV (4..7): all give the
// minimum-norm least-squares solution through the SVD.
enum LastMethod : int { kLastLU = 0, kLastLUPQ = 1, kLastQR = 2, kLastQRP = 3, kLastSVD = 4 };

// Arguments of the SVD last block (see bcr_last_svd.hip).
struct SvdArgs {
  int Nr, Nc, conditional;
  const double* A;          // Nr x Nc, ld Nr
  double *W, *Vj;           // scratch: max(Nr,Nc) x min(Nr,Nc), min^2
  double *U, *V, *S;        // Nr x k, Nc x k, k (k = min(Nr, Nc))
  int32_t* rank;
  int32_t* status;
};

// Arguments of the last-block solve: Hm is the factored Nr x Nc block
// (column-major, ld Nr); P = ipiv (LU, 1-based) / row swaps (LUPQ) / jpvt
// (QRP, 0-based); Q = column swaps (LUPQ); tau = reflectors (QR, QRP).
template <typename TX>
struct LastSolveArgs {
  int N, n, qr, qx, nr, nx, Nr, Nc, method, nrhs;
  const double* Hm; const int32_t* P; const int32_t* Q; const double* tau;
  TX* x; int64_t ldx;
  // SVD path: taken when method == kLastSVD, or when *status has bit 3 (the
  // second solver replaced a failed first factorization)
  const int32_t* status;
  const double *U, *V, *S; const int32_t* rank;
  const int32_t* skip = nullptr;   // batched: the system is done when *skip != 0
};

// Arguments of the FP64 block SpMV (pristine operator).
template <typename TA>
struct MvArgs {
  int N, n, qr, qx, nr, nx;
  const TA* D; const TA* E; const TA* B; const TA* C; const TA* H; const TA* Cq; const TA* F;
  const double* x;     // ncols
  const double* b;     // nrows (residual mode) or nullptr (y = A x)
  double* out;         // nrows
  double* Cpart;       // [N+1][nr] (per column with nrhs > 1)
  const int32_t* skip = nullptr;   // batched: the system is done when *skip != 0
  // several columns (residual mode): x, b, out advance by ldx, ldb, ldo
  int nrhs = 1;
  int64_t ldx = 0, ldb = 0, ldo = 0;
};

__host__ __device__ __forceinline__ size_t tsize(int n) { return size_t(2) * n * n + n; }
// block rows one wave of the residual SpMV takes: a block row uses
// g = max(n, nr) lanes (its n rows and its nr border partials), 64 / g rows
// per wave for g <= 32
__host__ __device__ __forceinline__ int mv_group(int n, int nr) { return n > nr ? n : (nr > 0 ? nr : 1); }
__host__ __device__ __forceinline__ int mv_rows_per_wave(int n, int nr) {
  const int g = mv_group(n, nr);
  return g <= 32 ? 64 / g : 1;
}
__host__ __device__ __forceinline__ size_t align16(size_t bytes) { return (bytes + 15) & ~size_t(15); }

template <typename T> __device__ __forceinline__ T dev_abs(T v) { return v < T(0) ? -v : v; }

// v where keep, +0 elsewhere, through an integer mask the compiler cannot see
// through.  A plain `keep ? v : 0` over a loaded v is turned into a branch
// around the load (the load sunk into it), which costs one full memory
// latency per element; with the mask every load is issued unconditionally
// (from a clamped, valid address) and all of them are in flight together.
__device__ __forceinline__ float keep_if(float v, bool keep) {
  int m = keep ? -1 : 0;
  asm volatile("" : "+v"(m));
  return __int_as_float(__float_as_int(v) & m);
}
__device__ __forceinline__ double keep_if(double v, bool keep) {
  int m = keep ? -1 : 0;
  asm volatile("" : "+v"(m));
  return __longlong_as_double(__double_as_longlong(v) & static_cast<long long>(m));
}

// element i of a device array p through a 32-bit byte offset (i * sizeof(T)
// < 2^32): with p the same in every lane of the wave the load takes the
// scalar-base form, one offset register and a 32-bit add per load instead of
// a 64-bit address, and a global access instead of a flat one (a pointer read
// from an argument record is otherwise treated as possibly pointing to LDS)
template <typename T>
__device__ __forceinline__ T gload(const T* p, unsigned i) {
  typedef const __attribute__((address_space(1))) char* gbytes;
  typedef const __attribute__((address_space(1))) T* gptr;
  return *(gptr)((gbytes)(p) + i * unsigned(sizeof(T)));
}

// batched refinement: a record whose system has converged (or stalled) is
// skipped by every kernel of the residual and correction launches
__device__ __forceinline__ bool skipped(const int32_t* f) { return f != nullptr && *f != 0; }

// pair p of level L (pure cyclic reduction over [0,N), cc:1088-1133)
// Batched launches: workgroup b serves entry e = entry[b] (one system at one
// level) as its local block b - first[e]; args[e] holds that system's
// arguments (pointers already offset to the system's slices).
template <typename Args>
struct BatchTable {
  const Args* args;
  const int32_t* entry;
  const int32_t* first;
};

__device__ __forceinline__ void pair_index(int p, int L, int N, int& j, int& jp, int& jpp) {
  j = L + 2 * L * p;
  jp = j - L;
  jpp = min(j + L, N);
}

// pair p of a launch: from the launch's pair table when it has one (the
// partitioned order of BABD_BorderedCR.cc:1072-1080 and 1181-1190: block
// ranges, then the chain of their end blocks), else pure cyclic reduction
template <typename Args>
__device__ __forceinline__ void pair_of(Args const& a, int p, int& j, int& jp, int& jpp);

// pair_of for a wave-uniform p: the block indices are moved into scalar
// registers, so addresses built from them keep a scalar base (the table
// lookup alone would leave them in vector registers, and every load from the
// pair's blocks would carry a 64-bit vector address)
template <typename Args>
__device__ __forceinline__ void pair_of_uniform(Args const& a, int p, int& j, int& jp, int& jpp) {
  pair_of(a, p, j, jp, jpp);
  j = __builtin_amdgcn_readfirstlane(j);
  jp = __builtin_amdgcn_readfirstlane(jp);
  jpp = __builtin_amdgcn_readfirstlane(jpp);
}

template <typename Args>
__device__ __forceinline__ void pair_of(Args const& a, int p, int& j, int& jp, int& jpp) {
  if (a.pairs != nullptr) {
    jp = a.pairs[3 * p];
    j = a.pairs[3 * p + 1];
    jpp = a.pairs[3 * p + 2];
  } else {
    pair_index(p, a.L, a.N, j, jp, jpp);
  }
}

}  // namespace bcr
