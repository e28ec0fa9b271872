// Pair elimination, LU, for block sizes up to 64: the columns of the
// augmented panel (R = 2n + nr rows, Wc = 3n + nx columns, the algebra of
// bcr_factor_generic.hip) are dealt out over the NW waves of a workgroup and
// its rows over the 64 lanes (row lane + 64 s, s < RS), all in registers.
// Wave w holds the T columns w + NW k (k < KT) and the right-hand columns
// n + w + NW k (k < KO).
//
// Per elimination step c, two barriers:
//  * wave 0 reads column c (published in LDS by its owner), finds the pivot:
//    a DPP max-reduction of |a_rc| over the rows at positions >= c among the
//    first 2n, ties to the smaller position (LAPACK idamax order), keeps the
//    row positions, and publishes the multipliers, the final column c and
//    the pivot lane;
//  * every wave then takes the pivot row's entry of each owned column from
//    the pivot lane by v_readlane (every wave holds every row of its
//    columns, so the pivot row is never published) and updates; the owner of
//    column c + 1 publishes it, updated, in the other LDS slot.  Finished T
//    columns (< c) are skipped.
// Rows are not moved: each is stored at its final position at the end.  The
// arithmetic is the generic kernel's (multiplier = a_rc * (1/pivot), fma
// update), so the factors agree with it bit for bit.
#include "bcr_launch.hh"
#include "bcr_batch_kernels.hh"
#include "bcr_wave.hh"

namespace bcr {
namespace {

// one element (r, s) of the augmented panel, from the source blocks
// (r < R, s < Wc); the pending-C rules are those of the generic kernel
template <typename TS, typename TW>
__device__ __forceinline__ TW panel_elem(FactorArgs<TS, TW> const& a, int r, int s, int j, int jp, TW* pend_cur,
                                         TW* pend_prv) {
  const int n = a.n, nr = a.nr, n2 = 2 * n;
  const size_t nn = size_t(n) * n, rn = size_t(nr) * n, bn = size_t(n) * a.nx;
  if (r < n) {                       // row of block jp: [E_jp | D_jp | 0 | B_jp]
    if (s < n) return TW(a.sE[jp * nn + r + size_t(s) * n]);
    if (s < n2) return TW(a.sD[jp * nn + r + size_t(s - n) * n]);
    if (s >= 3 * n) return TW(a.sB[jp * bn + r + size_t(s - 3 * n) * n]);
    return TW(0);
  }
  if (r < n2) {                      // row of block j: [D_j | 0 | E_j | B_j]
    const int rr = r - n;
    if (s < n) return TW(a.sD[j * nn + rr + size_t(s) * n]);
    if (s >= n2 && s < 3 * n) return TW(a.sE[j * nn + rr + size_t(s - n2) * n]);
    if (s >= 3 * n) return TW(a.sB[j * bn + rr + size_t(s - 3 * n) * n]);
    return TW(0);
  }
  const int ri = r - n2;             // border row: [C_j | C_jp | 0 | 0]
  if (s < n) {
    const size_t off = j * rn + ri + size_t(s) * nr;
    TW x = TW(a.sC[off]);
    if (!a.first) x += pend_prv[off] + pend_cur[off];
    return x;
  }
  if (s < n2) {
    const size_t off = jp * rn + ri + size_t(s - n) * nr;
    TW x = TW(a.sC[off]);
    if (!a.first) {
      x += pend_prv[off];
      pend_prv[off] = TW(0);   // single reader: clear for reuse
    }
    return x;
  }
  return TW(0);
}

// panel element (row at final position q, column s) to its factor block
template <typename TS, typename TW>
__device__ __forceinline__ void store_elem(FactorArgs<TS, TW> const& a, int q, int s, TW x, int j, int jp, int jpp,
                                           TW* Tj, TW* pend_cur) {
  const int n = a.n, nr = a.nr, n2 = 2 * n;
  const size_t nn = size_t(n) * n, rn = size_t(nr) * n, bn = size_t(n) * a.nx;
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
    else a.Fpart[size_t(j) * nr * a.nx + ri + size_t(s - 3 * n) * nr] = x;
  }
}

// the column updates of step c with the pivot row in row set sp, lane `who`;
// mc = the final values of column c (multipliers below the pivot)
template <int NW, int RS, int KT, int KO, typename TW>
__device__ __forceinline__ void update_columns(TW (&vt)[RS][KT], TW (&vo)[RS][KO], const TW (&lv)[RS],
                                               const TW (&mc)[RS], bool sp, int who, int w, int c, int n, int no,
                                               TW* next, int lane) {
  constexpr int S = RS > 1 ? 1 : 0;
#pragma unroll
  for (int k = 0; k < KT; ++k) {
    const int col = w + NW * k;
    if (col > c && col < n) {
      const TW u = wave::rl(sp ? vt[S][k] : vt[0][k], who);
#pragma unroll
      for (int s = 0; s < RS; ++s) vt[s][k] = fma(-lv[s], u, vt[s][k]);
      if (col == c + 1) {   // the next step's pivot column
#pragma unroll
        for (int s = 0; s < RS; ++s) next[lane + 64 * s] = vt[s][k];
      }
    } else if (col == c) {   // column c becomes the multipliers
#pragma unroll
      for (int s = 0; s < RS; ++s) vt[s][k] = mc[s];
    }
  }
  // right-hand columns in groups of four behind one guard: a padding column
  // (beyond 2n + nx) holds zeros and stays zero, so a group is updated whole
#pragma unroll
  for (int k0 = 0; k0 < KO; k0 += 4) {
    if (w + NW * k0 < no) {
#pragma unroll
      for (int k = k0; k < k0 + 4 && k < KO; ++k) {
        const TW u = wave::rl(sp ? vo[S][k] : vo[0][k], who);
#pragma unroll
        for (int s = 0; s < RS; ++s) vo[s][k] = fma(-lv[s], u, vo[s][k]);
      }
    }
    __builtin_amdgcn_sched_barrier(0);
  }
}

template <typename TS, typename TW, int NMAX, int NXMAX, int NW, int RS>
__device__ __forceinline__ void pair_lu_cols_body(FactorArgs<TS, TW> const& a, int p) {
  constexpr int KT = (NMAX + NW - 1) / NW;              // T columns per wave
  constexpr int KO = (2 * NMAX + NXMAX + NW - 1) / NW;  // D, E, B columns per wave
  __shared__ TW colbuf[2][RS * 64];
  __shared__ TW lbuf[RS * 64], mbuf[RS * 64];
  __shared__ int sbuf[2];
  __shared__ int posbuf[128];
  if (p >= a.npairs) return;   // uniform over the workgroup
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
  const int n = a.n, nr = a.nr, nx = a.nx, n2 = 2 * n;
  const int R = n2 + nr, no = n2 + nx;
  int j, jp, jpp;
  pair_of_uniform(a, p, j, jp, jpp);
  const size_t rn = size_t(nr) * n;
  TW* pend_cur = a.Cpend + size_t(a.slot_cur) * (a.N + 1) * rn;
  TW* pend_prv = a.Cpend + size_t(a.slot_cur ^ 1) * (a.N + 1) * rn;

  // ---- load --------------------------------------------------------------------
  TW vt[RS][KT], vo[RS][KO];
#pragma unroll
  for (int s = 0; s < RS; ++s) {
    const int r = lane + 64 * s;
#pragma unroll
    for (int k = 0; k < KT; ++k) {
      const int col = w + NW * k;
      vt[s][k] = (r < R && col < n) ? panel_elem(a, r, col, j, jp, pend_cur, pend_prv) : TW(0);
    }
#pragma unroll
    for (int k = 0; k < KO; ++k) {
      const int col = w + NW * k;
      vo[s][k] = (r < R && col < no) ? panel_elem(a, r, n + col, j, jp, pend_cur, pend_prv) : TW(0);
    }
  }
  if (w == 0) {
#pragma unroll
    for (int s = 0; s < RS; ++s) colbuf[0][lane + 64 * s] = vt[s][0];
  }
  // every wave has read its sources (some are rewritten at the end) and
  // column 0 is published
  __syncthreads();

  // ---- n elimination steps --------------------------------------------------------
  // Two barriers per step.  Wave 0 alone finds the pivot of column c, keeps
  // the row positions, forms the multipliers and publishes them (lbuf, with
  // zeros for the rows not updated), the final column c (mbuf) and the pivot
  // lane; then every wave updates its columns.  The search and the division
  // run once per step instead of once per wave: the step loop is bound by
  // VALU issue, not by its latency, when enough pairs share a CU.
  int pos0 = lane, pos1 = lane + 64;   // wave 0: positions of rows lane, lane + 64
  int myperm = 0;
  bool zero_pivot = false;
  for (int c = 0; c < n; ++c) {
    if (w == 0) {
      const TW* cur = colbuf[c & 1];
      TW col[RS];
#pragma unroll
      for (int s = 0; s < RS; ++s) col[s] = cur[lane + 64 * s];
      const bool c0 = lane < n2 && pos0 >= c;
      const bool c1 = RS > 1 && lane + 64 < n2 && pos1 >= c;
      unsigned h0, l0, h1 = 0, l1 = 0;
      wave::abs_key(col[0], h0, l0);
      if (RS > 1) wave::abs_key(col[RS > 1 ? 1 : 0], h1, l1);
      // the better of this lane's two candidates (ties to the smaller
      // position; bitwise operators: no short-circuit branches)
      const bool b1 = c1 & (!c0 | (h1 > h0) | ((h1 == h0) & ((l1 > l0) | ((l1 == l0) & (pos1 < pos0)))));
      bool zero;
      const int who = wave::argmax_pos<(sizeof(TW) > 4)>(c0 | c1, b1 ? h1 : h0, b1 ? l1 : l0, b1 ? pos1 : pos0,
                                                         lane, zero);
      const int P = __builtin_amdgcn_readlane(b1 ? pos1 : pos0, who);
      const int sp = __builtin_amdgcn_readlane(int(b1), who);
      zero_pivot = zero_pivot || zero;
      if (lane == c) myperm = P + 1;
      // the pivot row moves to position c, the row at c to the pivot's old place
      pos0 = pos0 == c ? P : pos0 == P ? c : pos0;
      pos1 = pos1 == c ? P : pos1 == P ? c : pos1;
      const TW pv = sp ? wave::rl(col[RS > 1 ? 1 : 0], who) : wave::rl(col[0], who);
      const TW rinv = pv != TW(0) ? TW(1) / pv : TW(0);
#pragma unroll
      for (int s = 0; s < RS; ++s) {   // rows below the pivot: multiplier, others 0
        const int r = lane + 64 * s;
        const bool upd = r < n2 ? (s == 0 ? pos0 > c : pos1 > c) : r < R;
        const TW l = col[s] * rinv;
        lbuf[lane + 64 * s] = upd ? l : TW(0);
        mbuf[lane + 64 * s] = upd ? l : col[s];
      }
      if (lane == 0) {
        sbuf[0] = who;
        sbuf[1] = sp;
      }
    }
    __syncthreads();   // pivot, multipliers and column c published
    TW lv[RS], mc[RS];
#pragma unroll
    for (int s = 0; s < RS; ++s) {
      lv[s] = lbuf[lane + 64 * s];
      mc[s] = mbuf[lane + 64 * s];
    }
    const int who = __builtin_amdgcn_readfirstlane(sbuf[0]);
    const int sp = __builtin_amdgcn_readfirstlane(sbuf[1]);
    TW* next = colbuf[(c + 1) & 1];
    // the column guards below compare compile-time k with these; opaque
    // copies keep the compiler from hoisting every guard out of the step loop
    // into its own SGPRs (which then spill and are reloaded each step)
    int w_ = w, n_ = n, no_ = no;
    asm volatile("" : "+s"(w_), "+s"(n_), "+s"(no_));
    update_columns<NW>(vt, vo, lv, mc, sp != 0, who, w_, c, n_, no_, next, lane);
    // column c + 1 published; every wave is done with this step's buffers
    __syncthreads();
  }
  if (w == 0) {
    if (zero_pivot && lane == 0) atomicOr(a.status, 1);
    posbuf[lane] = pos0;
    if (RS > 1) posbuf[64 + lane] = pos1;
  }
  __syncthreads();   // final row positions published

  // ---- store every row at its final position -----------------------------------
  TW* Tj = a.Tm + size_t(j) * tsize(n);
#pragma unroll
  for (int s = 0; s < RS; ++s) {
    const int r = lane + 64 * s;
    if (r >= R) continue;
    const int q = r < n2 ? posbuf[r] : r;
#pragma unroll
    for (int k = 0; k < KT; ++k) {
      const int col = w + NW * k;
      if (col < n) store_elem(a, q, col, vt[s][k], j, jp, jpp, Tj, pend_cur);
    }
#pragma unroll
    for (int k = 0; k < KO; ++k) {
      const int col = w + NW * k;
      if (col < no) store_elem(a, q, n + col, vo[s][k], j, jp, jpp, Tj, pend_cur);
    }
  }
  if (w == 0 && lane < n) a.perm[size_t(j) * n + lane] = myperm;
}

template <typename TS, typename TW, int NMAX, int NXMAX, int NW, int RS>
__global__ void __launch_bounds__(64 * NW) pair_lu_cols_kernel(FactorArgs<TS, TW> a) {
  pair_lu_cols_body<TS, TW, NMAX, NXMAX, NW, RS>(a, blockIdx.x);
}

template <typename TS, typename TW, int NMAX, int NXMAX, int NW, int RS>
__global__ void __launch_bounds__(64 * NW) pair_lu_cols_batched(BatchTable<FactorArgs<TS, TW>> t) {
  const int e = t.entry[blockIdx.x];
  pair_lu_cols_body<TS, TW, NMAX, NXMAX, NW, RS>(t.args[e], int(blockIdx.x) - t.first[e]);
}

// waves per workgroup: eight for n <= 48 (about 40 panel registers per lane
// in FP32, 80 in FP64: 7 and 4 waves per SIMD), 8 / 16 for n <= 64 (about 80
// registers); more waves per pair shorten each step's column loop, which is
// the step's critical path
template <typename TW> struct ColsWaves;
template <> struct ColsWaves<float> { static constexpr int k48 = 8, k64 = 8; };
template <> struct ColsWaves<double> { static constexpr int k48 = 8, k64 = 16; };

// 1: n <= 48, nx <= 4, 2n + nr <= 128 (two row sets); 2: n <= 64, nx <= 16,
// 2n + nr <= 192 (three row sets); 0: not covered
int cols_variant(int n, int nr, int nx) {
  if (n < 1) return 0;
  if (n <= 48 && nx <= 4 && 2 * n + nr <= 128) return 1;
  if (n <= 64 && nx <= 16 && 2 * n + nr <= 192) return 2;
  return 0;
}

}  // namespace

bool pair_lu_cols_fits(int n, int nr, int nx) { return cols_variant(n, nr, nx) != 0; }

template <typename TS, typename TW>
bool launch_pair_lu_cols(FactorArgs<TS, TW> const& a, hipStream_t st, hipError_t* err) {
  constexpr int W1 = ColsWaves<TW>::k48, W2 = ColsWaves<TW>::k64;
  switch (cols_variant(a.n, a.nr, a.nx)) {
    case 1:
      hipLaunchKernelGGL((pair_lu_cols_kernel<TS, TW, 48, 4, W1, 2>), dim3(a.npairs), dim3(64 * W1), 0, st, a);
      break;
    case 2:
      hipLaunchKernelGGL((pair_lu_cols_kernel<TS, TW, 64, 16, W2, 3>), dim3(a.npairs), dim3(64 * W2), 0, st, a);
      break;
    default:
      return false;
  }
  *err = hipGetLastError();
  return true;
}

template <typename TS, typename TW, int NMAX, int NXMAX, int NW, int RS>
void launch_cols_batched_w(BatchTable<FactorArgs<TS, TW>> const& t, int nblocks, hipStream_t st) {
  hipLaunchKernelGGL((pair_lu_cols_batched<TS, TW, NMAX, NXMAX, NW, RS>), dim3(nblocks), dim3(64 * NW), 0, st, t);
}

// waves per pair: 0 = the default of the element type; 4, 8 or 16 for
// n <= 48 (two row sets), 8 or 16 beyond (three row sets: four waves would
// hold 52 columns x 3 rows per lane)
template <typename TS, typename TW>
bool launch_pair_lu_cols_batched(BatchTable<FactorArgs<TS, TW>> const& t, int nblocks, int n, int nx, int nr,
                                 hipStream_t st, hipError_t* err, int waves) {
  constexpr int W1 = ColsWaves<TW>::k48, W2 = ColsWaves<TW>::k64;
  switch (cols_variant(n, nr, nx)) {
    case 1: {
      const int w = waves == 4 || waves == 8 || waves == 16 ? waves : W1;
      if (w == 4) launch_cols_batched_w<TS, TW, 48, 4, 4, 2>(t, nblocks, st);
      else if (w == 16) launch_cols_batched_w<TS, TW, 48, 4, 16, 2>(t, nblocks, st);
      else launch_cols_batched_w<TS, TW, 48, 4, 8, 2>(t, nblocks, st);
      break;
    }
    case 2: {
      const int w = waves == 8 || waves == 16 ? waves : W2;
      if (w == 16) launch_cols_batched_w<TS, TW, 64, 16, 16, 3>(t, nblocks, st);
      else launch_cols_batched_w<TS, TW, 64, 16, 8, 3>(t, nblocks, st);
      break;
    }
    default:
      return false;
  }
  *err = hipGetLastError();
  return true;
}

#define BCR_INST(TS, TW)                                                                                     \
  template bool launch_pair_lu_cols<TS, TW>(FactorArgs<TS, TW> const&, hipStream_t, hipError_t*);           \
  template bool launch_pair_lu_cols_batched<TS, TW>(BatchTable<FactorArgs<TS, TW>> const&, int, int, int, int, \
                                                    hipStream_t, hipError_t*, int);
BCR_INST(double, double)
BCR_INST(float, float)
BCR_INST(double, float)
#undef BCR_INST

}  // namespace bcr
