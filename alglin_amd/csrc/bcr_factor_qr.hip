// Pair elimination with Householder QR or column-pivoted QR (QRP), every
// n <= 64: the reference's buildT with geqrf / geqp3 and applyT with ormqr
// (BABD_BorderedCR.cc:866-891, 917-982) and the border update (cc:1107-1154).
//
// The augmented panel is the one of bcr_factor_generic.hip: R = 2n + nr rows
// [row jp; row j; border], columns [T = col j | col jp | col jpp | xb].  One
// 256-thread workgroup per pair, in two phases:
//
//  1. The T panel [E_jp; D_j; C_j] (R x n) in LDS: per column c the pivot
//     column (QRP: largest partial norm, dlaqp2 downdating), the reflector
//     (dlarfg), its application to the columns c+1..n-1 of T, and the
//     elimination of row c from the border rows (l = C[i][c] / R[c][c]);
//     one barrier per column (see the step loop).
//  2. The right-hand columns [col jp | col jpp | xb] (2n + nx of them), in
//     chunks of CW columns staged in LDS next to T: every column goes
//     through H_0 .. H_{n-1} and the border eliminations on its own, so a
//     group of G lanes (4 to 16: as many as keep every column of the chunk
//     in flight) owns a column for the whole sweep and no barrier is needed
//     inside it (the reflectors and border multipliers are read from T,
//     which phase 1 left in LDS).
//
// Each right-hand column sees the same operations in the same order as in a
// one-pass elimination of the whole panel; the dot products v^T x are summed
// in G interleaved partial sums (rows r = c + 1 + q + G k in lane q),
// combined so that the G lanes hold the same bits.  LDS columns have an odd
// stride (R | 1) so that a column per lane group reads conflict-free.  The
// chunk width follows the occupancy: the most workgroups per CU whose LDS
// share holds T and 16 columns, then the widest chunk that share allows.
// Results agree with LAPACK's to rounding; QRP's pivot order is LAPACK's
// except where two partial norms tie to rounding.
#include <algorithm>

#include "bcr_launch.hh"
#include "bcr_batch_kernels.hh"

namespace bcr {
namespace {

// lanes per column: a power of two from 4 to 16, the most that keeps every
// column of the step (or chunk) in flight at once
__device__ __forceinline__ int group_lanes(int cols) {
  const int per = 256 / (cols > 0 ? cols : 1);
  return per >= 16 ? 16 : per >= 8 ? 8 : 4;
}

// butterfly sum over the G lanes of a group: every lane ends with the same
// bits (each level adds the same two values, in either order)
template <typename T>
__device__ __forceinline__ T group_reduce(T d, int G) {
  for (int m = 1; m < G; m <<= 1) d += __shfl_xor(d, m);
  return d;
}

template <typename T>
__device__ __forceinline__ T qr_wave_sum(T v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// x <- H x, H = I - tau v v^T with v = (1, v[c+1..n2-1]) on rows c..n2-1,
// v[r] = u[r] * scal (the reflector before its scaling; the products round
// like the stored scaled vector would), by the G lanes of a group (lane q
// takes rows c + 1 + q + G k, two partial sums); returns the new x[c]
// (every lane of the group; lane 0 stores it).  (A fully unrolled form
// with every row's loads in flight at once measured 11-40 % slower: it
// loads the rows at and above c too, and its registers cost occupancy.)
template <typename T>
__device__ __forceinline__ T apply_reflector(const T* u, T scal, T* x, T tau, int c, int n2, int q, int G) {
  const T xc = x[c];
  T s0 = T(0), s1 = T(0);
  int r = c + 1 + q;
  for (; r + G < n2; r += 2 * G) {
    s0 = fma(u[r] * scal, x[r], s0);
    s1 = fma(u[r + G] * scal, x[r + G], s1);
  }
  if (r < n2) s0 = fma(u[r] * scal, x[r], s0);
  const T d = group_reduce(s0 + s1, G);
  const T w = (xc + d) * tau;
  for (r = c + 1 + q; r < n2; r += G) x[r] = fma(-(u[r] * scal), w, x[r]);
  const T nc = xc - w;
  if (q == 0) x[c] = nc;
  return nc;
}

// sum of squares of x[r0..r1-1] by the G lanes of a group, the same bits in
// every lane
template <typename T>
__device__ __forceinline__ T group_sumsq(const T* x, int r0, int r1, int q, int G) {
  T s = T(0);
  for (int r = r0 + q; r < r1; r += G) s = fma(x[r], x[r], s);
  return group_reduce(s, G);
}

constexpr int kLoad = 8;   // global loads a thread has in flight while staging

// dst(e) = src[e] (+ add1[e] + add2[e]) for e < cnt, element e at row e % rows,
// column e / rows of an LDS block with column stride ld; a thread issues
// kLoad loads before its first store.  clear: zero src-side slots after
// reading them (the pending border terms).
template <typename TS, typename TW>
__device__ __forceinline__ void stage(const TS* src, int cnt, int rows, TW* dst, int ld, int tid,
                                      TW* add1 = nullptr, const TW* add2 = nullptr, bool clear1 = false) {
  if (cnt <= 0) return;
  for (int base = 0; base < cnt; base += kLoad * 256) {
    TW v[kLoad];
#pragma unroll
    for (int k = 0; k < kLoad; ++k) {
      const int e = min(base + tid + k * 256, cnt - 1);
      v[k] = TW(src[e]);
    }
    if (add1) {
#pragma unroll
      for (int k = 0; k < kLoad; ++k) {
        const int e = min(base + tid + k * 256, cnt - 1);
        v[k] += add2 ? add1[e] + add2[e] : add1[e];
      }
    }
#pragma unroll
    for (int k = 0; k < kLoad; ++k) {
      const int e = base + tid + k * 256;
      if (e < cnt) {
        dst[(e % rows) + (e / rows) * ld] = v[k];
        if (clear1) add1[e] = TW(0);
      }
    }
  }
}

// rows r0..r1-1 of columns 0..cols-1 of an LDS block set to zero
template <typename TW>
__device__ __forceinline__ void zero_rows(TW* dst, int r0, int r1, int cols, int ld, int tid) {
  const int h = r1 - r0;
  for (int e = tid; e < h * cols; e += 256) dst[r0 + e % h + (e / h) * ld] = TW(0);
}

template <typename TS, typename TW, int METHOD>
__device__ __forceinline__ void pair_qr_body(FactorArgs<TS, TW> const& a, int cw, int p) {
  extern __shared__ __align__(16) unsigned char qr_smem[];
  __shared__ TW s_vn1[64], s_vn2[64], s_tau[64];
  __shared__ int s_map[64];
  if (p >= a.npairs) return;   // uniform over the workgroup
  int j, jp, jpp;
  pair_of_uniform(a, p, j, jp, jpp);
  const int n = a.n, nr = a.nr, nx = a.nx, n2 = 2 * n, R = n2 + nr;
  const int ld = R | 1;
  const int Wr = n2 + nx;   // right-hand columns
  const int tid = threadIdx.x, nth = blockDim.x;
  const int lane = tid & (kWave - 1);
  TW* P = reinterpret_cast<TW*>(qr_smem);   // T part, ld x n (physical columns)
  TW* X = P + size_t(ld) * n;                // right-hand chunk, ld x cw
  const size_t nn = size_t(n) * n, rn = size_t(nr) * n, bn = size_t(n) * nx;
  TW* pend_cur = a.Cpend + size_t(a.slot_cur) * (a.N + 1) * rn;
  TW* pend_prv = a.Cpend + size_t(a.slot_cur ^ 1) * (a.N + 1) * rn;
  const TS *const sD = a.sD, *const sE = a.sE, *const sB = a.sB, *const sC = a.sC;

  // ---- phase 1: T = [E_jp; D_j; C_j] ------------------------------------------
  stage(sE + jp * nn, n * n, n, P, ld, tid);
  stage(sD + j * nn, n * n, n, P + n, ld, tid);
  if (a.first) stage(sC + j * rn, nr * n, nr, P + n2, ld, tid);
  else stage(sC + j * rn, nr * n, nr, P + n2, ld, tid, pend_prv + j * rn, pend_cur + j * rn);
  __syncthreads();
  if (METHOD == kQRP) {
    for (int c = tid; c < n; c += nth) {
      TW s = 0;
      for (int r = 0; r < n2; ++r) {
        const TW v = P[r + c * ld];
        s += v * v;
      }
      s_vn1[c] = sqrt(s);
      s_vn2[c] = s_vn1[c];
    }
    __syncthreads();
  }
  // Step c works on logical column c.  QRP exchanges columns through a map
  // (logical -> physical), kept by every wave in a register (lane k holds
  // the physical column of logical column k) and updated the same way in
  // all waves, so a pivot exchange moves no data and needs no barrier.
  // Every wave forms the reflector itself (the same sums in the same
  // order); column c is scaled, given its diagonal and its border
  // multipliers in the next step, when no lane reads it.  One barrier per
  // step: the updated columns (and QRP norms) of step c for step c + 1.
  int colmap = lane;
  int pprev = 0, cprev = -1;
  TW sprev = TW(1), bprev = TW(0), iprev = TW(0);
  const TW tol3z = sqrt(TW(sizeof(TW) == 8 ? 2.220446049250313e-16 : 1.1920929e-07));
  for (int c = 0; c < n; ++c) {
    if (cprev >= 0) {   // finalize the previous column
      for (int r = cprev + 1 + tid; r < n2; r += nth) P[r + pprev * ld] *= sprev;
      if (tid == 0) P[cprev + pprev * ld] = bprev;
      for (int i = tid; i < nr; i += nth) P[n2 + i + pprev * ld] *= iprev;
    }
    if (METHOD == kQRP) {
      // the largest partial norm among logical columns c..n-1, ties to the
      // smaller logical index (LAPACK's idamax)
      TW best = lane >= c && lane < n ? s_vn1[colmap] : TW(-1);
      int bi = lane >= c && lane < n ? lane : 64;
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) {
        const TW ov = __shfl_xor(best, off);
        const int oi = __shfl_xor(bi, off);
        if (ov > best || (ov == best && oi < bi)) {
          best = ov;
          bi = oi;
        }
      }
      const int pvt = min(__builtin_amdgcn_readfirstlane(bi), n - 1);   // NaN norms: bi stays 64
      const int pc0 = __builtin_amdgcn_readlane(colmap, c), pp0 = __builtin_amdgcn_readlane(colmap, pvt);
      colmap = lane == c ? pp0 : lane == pvt ? pc0 : colmap;
    }
    const int pc = __builtin_amdgcn_readlane(colmap, c);   // physical column of step c
    // reflector of column c (dlarfg), in every wave
    TW ss = 0;
    for (int r = c + 1 + lane; r < n2; r += kWave) {
      const TW v = P[r + pc * ld];
      ss += v * v;
    }
    ss = qr_wave_sum(ss);
    ss = __shfl(ss, 0);
    const TW alpha = P[c + pc * ld];
    TW tau = 0, scal = 1, beta = alpha;
    if (ss != TW(0)) {
      const TW nrm = sqrt(alpha * alpha + ss);
      beta = alpha >= TW(0) ? -nrm : nrm;
      tau = (beta - alpha) / beta;
      scal = TW(1) / (alpha - beta);
    }
    const TW rinv = beta != TW(0) ? TW(1) / beta : TW(0);
    if (tid == 0) {
      s_tau[c] = tau;
      if (beta == TW(0)) atomicOr(a.status, 1);
    }
    // H_c on the logical columns c+1 .. n-1 (one per lane group of G
    // lanes), then their border rows lose row c, then (QRP) their norms
    const int G = group_lanes(n - c - 1);
    const int q = tid % G;
    const int s = c + 1 + tid / G;
    const int ps = __shfl(colmap, s < n ? s : 0);
    if (s < n) {
      TW* x = P + ps * ld;
      const TW nc = tau != TW(0) ? apply_reflector(P + pc * ld, scal, x, tau, c, n2, q, G) : x[c];
      for (int i = q; i < nr; i += G) x[n2 + i] = fma(-(P[n2 + i + pc * ld] * rinv), nc, x[n2 + i]);
      if (METHOD == kQRP) {
        // dlaqp2 norm downdating (LAWN 176)
        const TW v1 = s_vn1[ps];
        if (v1 != TW(0)) {
          TW t = dev_abs(nc) / v1;
          t = TW(1) - t * t;
          t = t < TW(0) ? TW(0) : t;
          const TW qq = v1 / s_vn2[ps];
          if (t * qq * qq <= tol3z) {
            const TW nv = sqrt(group_sumsq(x, c + 1, n2, q, G));
            if (q == 0) {
              s_vn1[ps] = nv;
              s_vn2[ps] = nv;
            }
          } else if (q == 0) {
            s_vn1[ps] = v1 * sqrt(t);
          }
        }
      }
    }
    pprev = pc;
    cprev = c;
    sprev = scal;
    bprev = beta;
    iprev = rinv;
    __syncthreads();
  }
  if (cprev >= 0) {
    for (int r = cprev + 1 + tid; r < n2; r += nth) P[r + pprev * ld] *= sprev;
    if (tid == 0) P[cprev + pprev * ld] = bprev;
    for (int i = tid; i < nr; i += nth) P[n2 + i + pprev * ld] *= iprev;
  }
  if (tid < n) s_map[tid] = colmap;   // wave 0's map (every wave holds the same)
  __syncthreads();

  // ---- phase 2: the right-hand columns, CW at a time ----------------------------
  // column g of [col jp | col jpp | xb]: region 0 (g < n) D_jp on top, C_jp in
  // the border rows; region 1 (g < 2n) E_j at the bottom; region 2 B_jp, B_j
  for (int g0 = 0; g0 < Wr; g0 += cw) {
    const int w = min(cw, Wr - g0);
    {
      const int a0 = g0, a1 = min(g0 + w, n);   // region 0 columns
      if (a1 > a0) {
        stage(sD + jp * nn + size_t(a0) * n, n * (a1 - a0), n, X, ld, tid);
        zero_rows(X, n, n2, a1 - a0, ld, tid);
        if (a.first) stage(sC + jp * rn + size_t(a0) * nr, nr * (a1 - a0), nr, X + n2, ld, tid);
        else
          stage(sC + jp * rn + size_t(a0) * nr, nr * (a1 - a0), nr, X + n2, ld, tid,
                pend_prv + jp * rn + size_t(a0) * nr, static_cast<const TW*>(nullptr), true);
      }
      const int b0 = max(g0, n), b1 = min(g0 + w, n2);   // region 1 columns
      if (b1 > b0) {
        TW* xb = X + (b0 - g0) * ld;
        zero_rows(xb, 0, n, b1 - b0, ld, tid);
        stage(sE + j * nn + size_t(b0 - n) * n, n * (b1 - b0), n, xb + n, ld, tid);
        zero_rows(xb, n2, R, b1 - b0, ld, tid);
      }
      const int c0 = max(g0, n2), c1 = g0 + w;   // region 2 columns
      if (c1 > c0) {
        TW* xc = X + (c0 - g0) * ld;
        stage(sB + jp * bn + size_t(c0 - n2) * n, n * (c1 - c0), n, xc, ld, tid);
        stage(sB + j * bn + size_t(c0 - n2) * n, n * (c1 - c0), n, xc + n, ld, tid);
        zero_rows(xc, n2, R, c1 - c0, ld, tid);
      }
    }
    __syncthreads();
    const int G = group_lanes(w), q = tid % G;
    for (int k = tid / G; k < w; k += 256 / G) {
      TW* x = X + k * ld;
      for (int c = 0; c < n; ++c) {
        const int pc = s_map[c];
        const TW tau = s_tau[c];
        const TW xc = tau != TW(0) ? apply_reflector(P + pc * ld, TW(1), x, tau, c, n2, q, G) : x[c];
        // the border rows lose row c (multipliers in T's border rows)
        for (int i = q; i < nr; i += G) x[n2 + i] = fma(-P[n2 + i + pc * ld], xc, x[n2 + i]);
        // the next step reads rows other lanes of the group wrote: keep the
        // LDS accesses of the steps in program order (one wave, in-order LDS)
        __builtin_amdgcn_wave_barrier();
      }
    }
    __syncthreads();
    for (int i = tid; i < R * w; i += nth) {
      const int r = i % R, k = i / R, g = g0 + k;
      const TW v = X[r + k * ld];
      if (g < n) {
        if (r < n) a.D[j * nn + r + size_t(g) * n] = v;
        else if (r < n2) a.D[jp * nn + (r - n) + size_t(g) * n] = v;
        else a.C[jp * rn + (r - n2) + size_t(g) * nr] = v;
      } else if (g < n2) {
        const int k1 = g - n;
        if (r < n) a.E[j * nn + r + size_t(k1) * n] = v;
        else if (r < n2) a.E[jp * nn + (r - n) + size_t(k1) * n] = v;
        else pend_cur[jpp * rn + (r - n2) + size_t(k1) * nr] += v;
      } else {
        const int k2 = g - n2;
        if (r < n) a.B[j * bn + r + size_t(k2) * n] = v;
        else if (r < n2) a.B[jp * bn + (r - n) + size_t(k2) * n] = v;
        else a.Fpart[size_t(j) * nr * nx + (r - n2) + size_t(k2) * nr] = v;
      }
    }
    __syncthreads();   // the chunk's slots are free again
  }

  // ---- T, tau, the column exchanges and C_j R^-1 (logical column order) -----------
  TW* Tj = a.Tm + size_t(j) * tsize(n);
  for (int i = tid; i < n2 * n; i += nth) {
    const int r = i % n2, c = i / n2;
    Tj[i] = P[r + s_map[c] * ld];
  }
  for (int c = tid; c < n; c += nth) Tj[size_t(n2) * n + c] = s_tau[c];
  for (int i = tid; i < nr * n; i += nth) {
    const int r = i % nr, c = i / nr;
    a.C[j * rn + i] = P[n2 + r + s_map[c] * ld];
  }
  if (tid == 0 && METHOD == kQRP) {
    // permutation (logical column c came from column s_map[c]) -> exchange
    // list (BABD_BorderedCR.hxx:265-275), worked out in place in s_map
    int32_t* Pj = a.perm + size_t(j) * n;
    for (int i = 0; i < n; ++i) {
      int jj = i;
      while (jj < n - 1 && s_map[jj] != i) ++jj;
      const int t = s_map[jj];
      s_map[jj] = s_map[i];
      s_map[i] = t;
      Pj[i] = jj;
    }
  }
}

template <typename TS, typename TW, int METHOD>
__global__ void __launch_bounds__(256) pair_qr_kernel(FactorArgs<TS, TW> a, int cw) {
  pair_qr_body<TS, TW, METHOD>(a, cw, blockIdx.x);
}

template <typename TS, typename TW, int METHOD>
__global__ void __launch_bounds__(256) pair_qr_batched(BatchTable<FactorArgs<TS, TW>> t, int cw) {
  const int e = t.entry[blockIdx.x];
  pair_qr_body<TS, TW, METHOD>(t.args[e], cw, int(blockIdx.x) - t.first[e]);
}

}  // namespace

bool pair_qr_plan(int n, int nr, int nx, size_t elem, size_t lds_limit, int* cw, size_t* bytes) {
  if (n < 1 || n > 64 || nr > 16) return false;
  const size_t ld = size_t(2 * n + nr) | 1;
  const int wr = 2 * n + nx;
  const int need = std::min(16, wr);   // columns of the smallest chunk
  // the most workgroups per CU (4 .. 1) whose share of the CU's LDS, less
  // the 2 KB of step tables, holds T and a chunk of `need` columns; then the
  // widest chunk that share allows
  for (int wgs = 4; wgs >= 1; --wgs) {
    const size_t share = std::min<size_t>(lds_limit, lds_limit / size_t(wgs)) - 2048;
    if (ld * size_t(n + need) * elem > share) continue;
    const int c = int(std::min<size_t>(size_t(wr), share / (ld * elem) - size_t(n)));
    *cw = c;
    *bytes = ld * size_t(n + c) * elem;
    return true;
  }
  return false;
}

template <typename TS, typename TW>
bool launch_pair_qr(int method, FactorArgs<TS, TW> const& a, size_t lds_limit, hipStream_t st, hipError_t* err) {
  int cw = 0;
  size_t bytes = 0;
  if (method == kLU || !pair_qr_plan(a.n, a.nr, a.nx, sizeof(TW), lds_limit, &cw, &bytes)) return false;
  const dim3 g(a.npairs), b(256);
  if (method == kQR) hipLaunchKernelGGL((pair_qr_kernel<TS, TW, kQR>), g, b, bytes, st, a, cw);
  else hipLaunchKernelGGL((pair_qr_kernel<TS, TW, kQRP>), g, b, bytes, st, a, cw);
  *err = hipGetLastError();
  return true;
}

template <typename TS, typename TW>
bool launch_pair_qr_batched(int method, BatchTable<FactorArgs<TS, TW>> const& t, int nblocks, int n, int nr, int nx,
                            size_t lds_limit, hipStream_t st, hipError_t* err) {
  int cw = 0;
  size_t bytes = 0;
  if (method == kLU || !pair_qr_plan(n, nr, nx, sizeof(TW), lds_limit, &cw, &bytes)) return false;
  const dim3 g(nblocks), b(256);
  if (method == kQR) hipLaunchKernelGGL((pair_qr_batched<TS, TW, kQR>), g, b, bytes, st, t, cw);
  else hipLaunchKernelGGL((pair_qr_batched<TS, TW, kQRP>), g, b, bytes, st, t, cw);
  *err = hipGetLastError();
  return true;
}

template bool launch_pair_qr<double, double>(int, FactorArgs<double, double> const&, size_t, hipStream_t, hipError_t*);
template bool launch_pair_qr<float, float>(int, FactorArgs<float, float> const&, size_t, hipStream_t, hipError_t*);
template bool launch_pair_qr<double, float>(int, FactorArgs<double, float> const&, size_t, hipStream_t, hipError_t*);
template bool launch_pair_qr_batched<double, double>(int, BatchTable<FactorArgs<double, double>> const&, int, int,
                                                     int, int, size_t, hipStream_t, hipError_t*);
template bool launch_pair_qr_batched<float, float>(int, BatchTable<FactorArgs<float, float>> const&, int, int, int,
                                                   int, size_t, hipStream_t, hipError_t*);
template bool launch_pair_qr_batched<double, float>(int, BatchTable<FactorArgs<double, float>> const&, int, int, int,
                                                    int, size_t, hipStream_t, hipError_t*);

}  // namespace bcr
