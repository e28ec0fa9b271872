// Pair elimination, LU, for block sizes up to 64 (the augmented panel of
// bcr_factor_generic.hip: R = 2n + nr rows, Wc = 3n + nx columns), with the
// panel held in registers instead of LDS.
//
// 512 threads on a 32 x 16 grid; thread (tr, tc) owns panel rows tr + 32 i and
// columns tc + 16 k (cyclic, so the shrinking trailing block stays spread
// over every thread), at most 5 x 13 values.  The panel is loaded straight
// from the source blocks and stored straight to the factor blocks.  Rows are
// not moved: every row keeps a position (LDS), the pivot of column c is the
// first maximum of |a_rc| over the rows at positions >= c among the first 2n
// rows (ties to the smaller position, LAPACK idamax order), and rows are
// written to their final positions at the end.  Per step: the owners of
// column c publish it (LDS), one wave picks the pivot, the owners of the
// pivot row publish its trailing part, and every thread updates its rows
// below the pivot -- two LDS buffers and three barriers per step instead of
// re-reading and re-writing the whole trailing block in LDS.  The arithmetic
// is the generic kernel's (multiplier = a_rc * (1/pivot), fma update), so the
// factors agree with it bit for bit.
#include "bcr_launch.hh"
#include "bcr_batch_kernels.hh"

namespace bcr {
namespace {

constexpr int TR = 32, TC = 16;
// tile depths: MR x MC values per thread cover R <= 32 MR rows and Wc <= 16 MC
// columns; (5, 13) serves every n <= 64, (3, 7) the n <= 32 panels with fewer
// registers (more workgroups per CU)
constexpr int kBigMR = 5, kBigMC = 13, kSmallMR = 3, kSmallMC = 7;

template <typename TS, typename TW, int MR, int MC>
__device__ __forceinline__ void pair_lu_tiled_body(FactorArgs<TS, TW> a, int bid) {
  constexpr int RMAX = TR * MR, WMAX = TC * MC;
  __shared__ TW colbuf2[2][RMAX];   // column c in slot c & 1: a thread may still read step c's
                                    // while the owners of column c+1 write theirs
  __shared__ TW rowbuf[WMAX];
  __shared__ int spos[RMAX], srow[RMAX];
  __shared__ int s_piv, s_pivpos;
  __shared__ int s_ipiv[64];
  const int p = bid;
  if (p >= a.npairs) return;   // uniform over the workgroup
  int j, jp, jpp;
  pair_of_uniform(a, p, j, jp, jpp);
  const int n = a.n, nr = a.nr, nx = a.nx, n2 = 2 * n;
  const int R = n2 + nr, Wc = 3 * n + nx;
  const int tid = threadIdx.x, tr = tid % TR, tc = tid / TR;
  const int lane = tid & 63, wid = tid >> 6;
  const size_t nn = size_t(n) * n, rn = size_t(nr) * n, bn = size_t(n) * nx;
  TW* pend_cur = a.Cpend + size_t(a.slot_cur) * (a.N + 1) * rn;
  TW* pend_prv = a.Cpend + size_t(a.slot_cur ^ 1) * (a.N + 1) * rn;

  // ---- load the panel into the register tiles ---------------------------------
  TW v[MR][MC];
#pragma unroll
  for (int i = 0; i < MR; ++i) {
    const int r = tr + TR * i;
#pragma unroll
    for (int k = 0; k < MC; ++k) {
      const int s = tc + TC * k;
      TW x = TW(0);
      if (r < R && s < Wc) {
        if (r < n) {                       // row of block jp: [E_jp | D_jp | 0 | B_jp]
          if (s < n) x = TW(a.sE[jp * nn + r + size_t(s) * n]);
          else if (s < n2) x = TW(a.sD[jp * nn + r + size_t(s - n) * n]);
          else if (s >= 3 * n) x = TW(a.sB[jp * bn + r + size_t(s - 3 * n) * n]);
        } else if (r < n2) {               // row of block j: [D_j | 0 | E_j | B_j]
          const int rr = r - n;
          if (s < n) x = TW(a.sD[j * nn + rr + size_t(s) * n]);
          else if (s >= n2 && s < 3 * n) x = TW(a.sE[j * nn + rr + size_t(s - n2) * n]);
          else if (s >= 3 * n) x = TW(a.sB[j * bn + rr + size_t(s - 3 * n) * n]);
        } else {                           // border row: [C_j | C_jp | 0 | 0]
          const int ri = r - n2;
          if (s < n) {
            const size_t off = j * rn + ri + size_t(s) * nr;
            x = TW(a.sC[off]);
            if (!a.first) x += pend_prv[off] + pend_cur[off];
          } else if (s < n2) {
            const size_t off = jp * rn + ri + size_t(s - n) * nr;
            x = TW(a.sC[off]);
            if (!a.first) {
              x += pend_prv[off];
              pend_prv[off] = TW(0);   // single reader: clear for reuse
            }
          }
        }
      }
      v[i][k] = x;
    }
  }
  for (int r = tid; r < R; r += blockDim.x) { spos[r] = r; srow[r] = r; }
  // every thread has read its sources (some are rewritten below)
  __syncthreads();

  // ---- n elimination steps --------------------------------------------------
  for (int c = 0; c < n; ++c) {
    const int kc = c / TC;
    TW* colbuf = colbuf2[c & 1];
    if (tc == c % TC) {
#pragma unroll
      for (int i = 0; i < MR; ++i) {
        const int r = tr + TR * i;
        TW x = TW(0);
#pragma unroll
        for (int k = 0; k < MC; ++k) x = (k == kc) ? v[i][k] : x;
        if (r < R) colbuf[r] = x;
      }
    }
    __syncthreads();
    if (wid == 0) {
      TW best = TW(-1);
      int bpos = 0x7fffffff, brow = 0;
      for (int r = lane; r < n2; r += 64) {
        const int q = spos[r];
        if (q < c) continue;
        const TW x = dev_abs(colbuf[r]);
        if (x > best || (x == best && q < bpos)) { best = x; bpos = q; brow = r; }
      }
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) {
        const TW ob = __shfl_xor(best, off);
        const int op = __shfl_xor(bpos, off), orow = __shfl_xor(brow, off);
        if (ob > best || (ob == best && op < bpos)) { best = ob; bpos = op; brow = orow; }
      }
      if (lane == 0) {
        s_piv = brow;
        s_pivpos = bpos;
        s_ipiv[c] = bpos + 1;
        if (best == TW(0)) atomicOr(a.status, 1);
        // positions: the pivot row moves to c, the row at c to the pivot's old place
        const int q = srow[c];
        spos[q] = bpos;
        srow[bpos] = q;
        spos[brow] = c;
        srow[c] = brow;
      }
    }
    __syncthreads();
    // the pivot row's trailing part, zero elsewhere (columns <= c and the
    // padding), so the update below needs no per-element test
    const int prow = s_piv;
    if (tr == prow % TR) {
      const int ip = prow / TR;
#pragma unroll
      for (int k = 0; k < MC; ++k) {
        const int s = tc + TC * k;
        TW x = TW(0);
#pragma unroll
        for (int i = 0; i < MR; ++i) x = (i == ip) ? v[i][k] : x;
        rowbuf[s] = (s > c && s < Wc) ? x : TW(0);
      }
    }
    __syncthreads();
    const TW pv = colbuf[prow];
    const TW rinv = pv != TW(0) ? TW(1) / pv : TW(0);
    TW lv[MR], rb[MC];
#pragma unroll
    for (int i = 0; i < MR; ++i) {   // multiplier of each row, zero for rows not updated
      const int r = tr + TR * i;
      lv[i] = (r < R && spos[r] > c) ? colbuf[r] * rinv : TW(0);
    }
#pragma unroll
    for (int k = 0; k < MC; ++k) rb[k] = rowbuf[tc + TC * k];
#pragma unroll
    for (int i = 0; i < MR; ++i)
#pragma unroll
      for (int k = 0; k < MC; ++k) v[i][k] = fma(-lv[i], rb[k], v[i][k]);
    if (tc == c % TC) {   // column c becomes the multipliers
#pragma unroll
      for (int i = 0; i < MR; ++i) {
        const int r = tr + TR * i;
        const bool u = r < R && spos[r] > c;
#pragma unroll
        for (int k = 0; k < MC; ++k) v[i][k] = (u && k == kc) ? lv[i] : v[i][k];
      }
    }
    // rowbuf is rewritten next step only after its second barrier
  }
  __syncthreads();

  // ---- store every row at its final position ---------------------------------
  TW* Tj = a.Tm + size_t(j) * tsize(n);
#pragma unroll
  for (int i = 0; i < MR; ++i) {
    const int r = tr + TR * i;
    if (r >= R) continue;
    const int q = spos[r];
#pragma unroll
    for (int k = 0; k < MC; ++k) {
      const int s = tc + TC * k;
      if (s >= Wc) continue;
      const TW x = v[i][k];
      if (q < n2) {
        const bool piv = q < n;               // pivot rows -> block j, reduced rows -> block jp
        const int rp = piv ? q : q - n;
        const size_t blk = piv ? size_t(j) : size_t(jp);
        if (s < n) Tj[q + size_t(s) * n2] = x;
        else if (s < n2) a.D[blk * nn + rp + size_t(s - n) * n] = x;
        else if (s < 3 * n) a.E[blk * nn + rp + size_t(s - n2) * n] = x;
        else a.B[blk * bn + rp + size_t(s - 3 * n) * n] = x;
      } else {
        const int ri = q - n2;
        if (s < n) a.C[j * rn + ri + size_t(s) * nr] = x;
        else if (s < n2) a.C[jp * rn + ri + size_t(s - n) * nr] = x;
        else if (s < 3 * n) pend_cur[jpp * rn + ri + size_t(s - n2) * nr] += x;
        else a.Fpart[size_t(j) * nr * nx + ri + size_t(s - 3 * n) * nr] = x;
      }
    }
  }
  for (int c = tid; c < n; c += blockDim.x) a.perm[size_t(j) * n + c] = s_ipiv[c];
}

template <typename TS, typename TW, int MR, int MC>
__global__ void __launch_bounds__(512) pair_lu_tiled_kernel(FactorArgs<TS, TW> a) {
  pair_lu_tiled_body<TS, TW, MR, MC>(a, blockIdx.x);
}

template <typename TS, typename TW, int MR, int MC>
__global__ void __launch_bounds__(512) pair_lu_tiled_batched(BatchTable<FactorArgs<TS, TW>> t) {
  const int e = t.entry[blockIdx.x];
  pair_lu_tiled_body<TS, TW, MR, MC>(t.args[e], int(blockIdx.x) - t.first[e]);
}

bool fits(int n, int nr, int nx, int mr, int mc) { return n <= 64 && 2 * n + nr <= TR * mr && 3 * n + nx <= TC * mc; }

}  // namespace

bool pair_lu_tiled_fits(int n, int nr, int nx) { return fits(n, nr, nx, kBigMR, kBigMC); }

template <typename TS, typename TW>
bool launch_pair_lu_tiled(FactorArgs<TS, TW> const& a, hipStream_t st, hipError_t* err) {
  if (fits(a.n, a.nr, a.nx, kSmallMR, kSmallMC))
    hipLaunchKernelGGL((pair_lu_tiled_kernel<TS, TW, kSmallMR, kSmallMC>), dim3(a.npairs), dim3(512), 0, st, a);
  else if (fits(a.n, a.nr, a.nx, kBigMR, kBigMC))
    hipLaunchKernelGGL((pair_lu_tiled_kernel<TS, TW, kBigMR, kBigMC>), dim3(a.npairs), dim3(512), 0, st, a);
  else
    return false;
  *err = hipGetLastError();
  return true;
}

template <typename TS, typename TW>
bool launch_pair_lu_tiled_batched(BatchTable<FactorArgs<TS, TW>> const& t, int nblocks, int n, int nx, int nr,
                                  hipStream_t st, hipError_t* err) {
  if (fits(n, nr, nx, kSmallMR, kSmallMC))
    hipLaunchKernelGGL((pair_lu_tiled_batched<TS, TW, kSmallMR, kSmallMC>), dim3(nblocks), dim3(512), 0, st, t);
  else if (fits(n, nr, nx, kBigMR, kBigMC))
    hipLaunchKernelGGL((pair_lu_tiled_batched<TS, TW, kBigMR, kBigMC>), dim3(nblocks), dim3(512), 0, st, t);
  else
    return false;
  *err = hipGetLastError();
  return true;
}

template bool launch_pair_lu_tiled<double, double>(FactorArgs<double, double> const&, hipStream_t, hipError_t*);
template bool launch_pair_lu_tiled<float, float>(FactorArgs<float, float> const&, hipStream_t, hipError_t*);
template bool launch_pair_lu_tiled<double, float>(FactorArgs<double, float> const&, hipStream_t, hipError_t*);
template bool launch_pair_lu_tiled_batched<double, double>(BatchTable<FactorArgs<double, double>> const&, int, int, int,
                                                           int, hipStream_t, hipError_t*);
template bool launch_pair_lu_tiled_batched<float, float>(BatchTable<FactorArgs<float, float>> const&, int, int, int,
                                                         int, hipStream_t, hipError_t*);
template bool launch_pair_lu_tiled_batched<double, float>(BatchTable<FactorArgs<double, float>> const&, int, int, int,
                                                          int, hipStream_t, hipError_t*);

}  // namespace bcr
