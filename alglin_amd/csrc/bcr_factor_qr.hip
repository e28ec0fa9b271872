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
//     elimination of row c from the border rows (l = C[i][c] / R[c][c]).
//  2. The right-hand columns [col jp | col jpp | xb] (2n + nx of them), in
//     chunks of CW columns staged in LDS next to T: every column goes
//     through H_0 .. H_{n-1} and the border eliminations on its own, so a
//     group of four lanes owns a column for the whole sweep and no barrier
//     is needed inside it (the reflectors and border multipliers are read
//     from T, which phase 1 left in LDS).
//
// Each right-hand column sees the same operations in the same order as in a
// one-pass elimination of the whole panel; the dot products v^T x are summed
// in four interleaved partial sums (rows r = c + 1 + q + 4k in lane q),
// combined so that the four lanes hold the same bits.  LDS columns have an
// odd stride (R | 1) so that a column per lane group reads conflict-free.
// Results agree with LAPACK's to rounding; QRP's pivot order is LAPACK's
// except where two partial norms tie to rounding.
#include <algorithm>

#include "bcr_launch.hh"
#include "bcr_batch_kernels.hh"

namespace bcr {
namespace {

constexpr int kGroup = 4;                  // lanes per column
constexpr int kGroups = 256 / kGroup;      // columns in flight

template <typename T>
__device__ __forceinline__ T qr_wave_sum(T v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// v^T x over rows c+1 .. n2-1 in four partial sums, the same bits in the
// four lanes of the group
template <typename T>
__device__ __forceinline__ T group_dot(const T* v, const T* x, int c, int n2, int q) {
  T s0 = T(0), s1 = T(0);
  int r = c + 1 + q;
  for (; r + kGroup < n2; r += 2 * kGroup) {
    s0 = fma(v[r], x[r], s0);
    s1 = fma(v[r + kGroup], x[r + kGroup], s1);
  }
  if (r < n2) s0 = fma(v[r], x[r], s0);
  T p = s0 + s1;
  p += __shfl_xor(p, 1);
  p += __shfl_xor(p, 2);
  return p;
}

// x <- H x, H = I - tau v v^T with v = (1, v[c+1..n2-1]) on rows c..n2-1;
// returns the new x[c] (all four lanes; lane 0 stores it)
template <typename T>
__device__ __forceinline__ T apply_reflector(const T* v, T* x, T tau, int c, int n2, int q) {
  const T xc = x[c];
  const T w = (xc + group_dot(v, x, c, n2, q)) * tau;
  for (int r = c + 1 + q; r < n2; r += kGroup) x[r] = fma(-v[r], w, x[r]);
  const T nc = xc - w;
  if (q == 0) x[c] = nc;
  return nc;
}

template <typename TS, typename TW, int METHOD>
__device__ __forceinline__ void pair_qr_body(FactorArgs<TS, TW> const& a, int cw, int p) {
  extern __shared__ __align__(16) unsigned char qr_smem[];
  __shared__ TW s_vn1[64], s_vn2[64], s_tau[64], s_lr[16];
  __shared__ TW s_red[2];
  __shared__ int s_piv;
  __shared__ int s_jpvt[64];
  if (p >= a.npairs) return;   // uniform over the workgroup
  int j, jp, jpp;
  pair_of(a, p, j, jp, jpp);
  const int n = a.n, nr = a.nr, nx = a.nx, n2 = 2 * n, R = n2 + nr;
  const int ld = R | 1;
  const int Wr = n2 + nx;   // right-hand columns
  const int tid = threadIdx.x, nth = blockDim.x;
  const int lane = tid & (kWave - 1), wid = tid / kWave;
  const int grp = tid / kGroup, q = tid % kGroup;
  TW* P = reinterpret_cast<TW*>(qr_smem);   // T part, ld x n
  TW* X = P + size_t(ld) * n;                // right-hand chunk, ld x cw
  const size_t nn = size_t(n) * n, rn = size_t(nr) * n, bn = size_t(n) * nx;
  TW* pend_cur = a.Cpend + size_t(a.slot_cur) * (a.N + 1) * rn;
  TW* pend_prv = a.Cpend + size_t(a.slot_cur ^ 1) * (a.N + 1) * rn;

  // ---- phase 1: T = [E_jp; D_j; C_j] ------------------------------------------
  for (int i = tid; i < n * n; i += nth) {
    const int r = i % n, c = i / n;
    P[r + c * ld] = TW(a.sE[jp * nn + i]);
    P[n + r + c * ld] = TW(a.sD[j * nn + i]);
  }
  for (int i = tid; i < nr * n; i += nth) {
    const int r = i % nr, c = i / nr;
    TW cj = TW(a.sC[j * rn + i]);
    if (!a.first) cj += pend_prv[j * rn + i] + pend_cur[j * rn + i];
    P[n2 + r + c * ld] = cj;
  }
  if (METHOD == kQRP) {
    __syncthreads();
    for (int c = tid; c < n; c += nth) {
      TW s = 0;
      for (int r = 0; r < n2; ++r) {
        const TW v = P[r + c * ld];
        s += v * v;
      }
      s_vn1[c] = sqrt(s);
      s_vn2[c] = s_vn1[c];
      s_jpvt[c] = c;
    }
  }
  __syncthreads();
  for (int c = 0; c < n; ++c) {
    if (METHOD == kQRP) {
      if (wid == 0) {
        TW best = TW(-1);
        int bi = c;   // valid also when every norm is NaN
        for (int s = c + lane; s < n; s += kWave)
          if (s_vn1[s] > best) {
            best = s_vn1[s];
            bi = s;
          }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {   // ties keep the smaller index
          const TW ov = __shfl_xor(best, off);
          const int oi = __shfl_xor(bi, off);
          if (ov > best || (ov == best && oi < bi)) {
            best = ov;
            bi = oi;
          }
        }
        if (lane == 0) s_piv = bi;
      }
      __syncthreads();
      const int pvt = s_piv;
      if (pvt != c) {
        for (int r = tid; r < R; r += nth) {   // whole column incl. the border rows
          const TW t = P[r + c * ld];
          P[r + c * ld] = P[r + pvt * ld];
          P[r + pvt * ld] = t;
        }
        if (tid == 0) {
          const int t = s_jpvt[c];
          s_jpvt[c] = s_jpvt[pvt];
          s_jpvt[pvt] = t;
          s_vn1[pvt] = s_vn1[c];
          s_vn2[pvt] = s_vn2[c];
        }
      }
      __syncthreads();
    }
    // reflector of column c (dlarfg)
    if (wid == 0) {
      TW ss = 0;
      for (int r = c + 1 + lane; r < n2; r += kWave) {
        const TW v = P[r + c * ld];
        ss += v * v;
      }
      ss = qr_wave_sum(ss);
      if (lane == 0) {
        const TW alpha = P[c + c * ld];
        TW tau = 0, scal = 1, beta = alpha;
        if (ss != TW(0)) {
          const TW nrm = sqrt(alpha * alpha + ss);
          beta = alpha >= TW(0) ? -nrm : nrm;
          tau = (beta - alpha) / beta;
          scal = TW(1) / (alpha - beta);
        }
        s_tau[c] = tau;
        s_red[0] = scal;
        s_red[1] = beta;
      }
    }
    __syncthreads();
    {
      const TW scal = s_red[0];
      for (int r = c + 1 + tid; r < n2; r += nth) P[r + c * ld] *= scal;
      if (tid == 0) P[c + c * ld] = s_red[1];
    }
    __syncthreads();
    // H_c on the T columns c+1 .. n-1, one lane group per column
    const TW tau = s_tau[c];
    if (tau != TW(0))
      for (int s = c + 1 + grp; s < n; s += kGroups) apply_reflector(P + c * ld, P + s * ld, tau, c, n2, q);
    __syncthreads();
    // row c is final: its multipliers for the border rows
    {
      const TW rc = P[c + c * ld];
      if (rc == TW(0) && tid == 0) atomicOr(a.status, 1);
      const TW rinv = rc != TW(0) ? TW(1) / rc : TW(0);
      for (int i = tid; i < nr; i += nth) {
        const TW l = P[n2 + i + c * ld] * rinv;
        P[n2 + i + c * ld] = l;
        s_lr[i] = l;
      }
    }
    __syncthreads();
    for (int i = tid; i < nr * (n - c - 1); i += nth) {
      const int r = i % nr, s = c + 1 + i / nr;
      P[n2 + r + s * ld] -= s_lr[r] * P[c + s * ld];
    }
    if (METHOD == kQRP) {
      // dlaqp2 norm downdating (LAWN 176)
      const TW tol3z = sqrt(TW(sizeof(TW) == 8 ? 2.220446049250313e-16 : 1.1920929e-07));
      for (int s = c + 1 + tid; s < n; s += nth) {
        if (s_vn1[s] != TW(0)) {
          TW t = dev_abs(P[c + s * ld]) / s_vn1[s];
          t = TW(1) - t * t;
          t = t < TW(0) ? TW(0) : t;
          const TW qq = s_vn1[s] / s_vn2[s];
          if (t * qq * qq <= tol3z) {
            TW ssum = 0;
            for (int r = c + 1; r < n2; ++r) {
              const TW v = P[r + s * ld];
              ssum += v * v;
            }
            s_vn1[s] = sqrt(ssum);
            s_vn2[s] = s_vn1[s];
          } else {
            s_vn1[s] *= sqrt(t);
          }
        }
      }
    }
    __syncthreads();
  }

  // ---- phase 2: the right-hand columns, CW at a time ----------------------------
  // column g of [col jp | col jpp | xb]: region 0 (g < n) D_jp on top, C_jp in
  // the border rows; region 1 (g < 2n) E_j at the bottom; region 2 B_jp, B_j
  for (int g0 = 0; g0 < Wr; g0 += cw) {
    const int w = min(cw, Wr - g0);
    for (int i = tid; i < R * w; i += nth) {
      const int r = i % R, k = i / R, g = g0 + k;
      TW v = TW(0);
      if (g < n) {
        if (r < n) v = TW(a.sD[jp * nn + r + size_t(g) * n]);
        else if (r >= n2) {
          const size_t e = jp * rn + (r - n2) + size_t(g) * nr;
          v = TW(a.sC[e]);
          if (!a.first) {
            v += pend_prv[e];
            pend_prv[e] = TW(0);
          }
        }
      } else if (g < n2) {
        if (r >= n && r < n2) v = TW(a.sE[j * nn + (r - n) + size_t(g - n) * n]);
      } else {
        const int k2 = g - n2;
        if (r < n) v = TW(a.sB[jp * bn + r + size_t(k2) * n]);
        else if (r < n2) v = TW(a.sB[j * bn + (r - n) + size_t(k2) * n]);
      }
      X[r + k * ld] = v;
    }
    __syncthreads();
    for (int k = grp; k < w; k += kGroups) {
      TW* x = X + k * ld;
      for (int c = 0; c < n; ++c) {
        const TW tau = s_tau[c];
        const TW xc = tau != TW(0) ? apply_reflector(P + c * ld, x, tau, c, n2, q) : x[c];
        // the border rows lose row c (multipliers in T's border rows)
        for (int i = q; i < nr; i += kGroup) x[n2 + i] = fma(-P[n2 + i + c * ld], xc, x[n2 + i]);
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

  // ---- T, tau, the column exchanges and C_j R^-1 ----------------------------------
  TW* Tj = a.Tm + size_t(j) * tsize(n);
  for (int i = tid; i < n2 * n; i += nth) {
    const int r = i % n2, c = i / n2;
    Tj[i] = P[r + c * ld];
  }
  for (int c = tid; c < n; c += nth) Tj[size_t(n2) * n + c] = s_tau[c];
  for (int i = tid; i < nr * n; i += nth) {
    const int r = i % nr, c = i / nr;
    a.C[j * rn + i] = P[n2 + r + c * ld];
  }
  if (tid == 0 && METHOD == kQRP) {
    // permutation -> exchange list (BABD_BorderedCR.hxx:265-275), worked
    // out in place in s_jpvt (0-based entries)
    int32_t* Pj = a.perm + size_t(j) * n;
    for (int i = 0; i < n; ++i) {
      int jj = i;
      while (jj < n && s_jpvt[jj] != i) ++jj;
      const int t = s_jpvt[jj];
      s_jpvt[jj] = s_jpvt[i];
      s_jpvt[i] = t;
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
  // aim at about 40 KB (four workgroups per CU), at least 16 columns a chunk
  const long budget = long(std::min<size_t>(lds_limit, 40 * 1024) / (ld * elem)) - n;
  int c = int(std::max<long>(16, budget));
  c = std::min(c, wr);
  const size_t b = ld * size_t(n + c) * elem;
  if (b > lds_limit) return false;
  *cw = c;
  *bytes = b;
  return true;
}

template <typename TS, typename TW>
bool launch_pair_qr(int method, FactorArgs<TS, TW> const& a, size_t lds_limit, hipStream_t st, hipError_t* err) {
  int cw = 0;
  size_t bytes = 0;
  if (method == kLU || !pair_qr_plan(a.n, a.nr, a.nx, sizeof(TW), lds_limit, &cw, &bytes)) return false;
  if (method == kQR)
    hipLaunchKernelGGL((pair_qr_kernel<TS, TW, kQR>), dim3(a.npairs), dim3(256), bytes, st, a, cw);
  else
    hipLaunchKernelGGL((pair_qr_kernel<TS, TW, kQRP>), dim3(a.npairs), dim3(256), bytes, st, a, cw);
  *err = hipGetLastError();
  return true;
}

template <typename TS, typename TW>
bool launch_pair_qr_batched(int method, BatchTable<FactorArgs<TS, TW>> const& t, int nblocks, int n, int nr, int nx,
                            size_t lds_limit, hipStream_t st, hipError_t* err) {
  int cw = 0;
  size_t bytes = 0;
  if (method == kLU || !pair_qr_plan(n, nr, nx, sizeof(TW), lds_limit, &cw, &bytes)) return false;
  if (method == kQR)
    hipLaunchKernelGGL((pair_qr_batched<TS, TW, kQR>), dim3(nblocks), dim3(256), bytes, st, t, cw);
  else
    hipLaunchKernelGGL((pair_qr_batched<TS, TW, kQRP>), dim3(nblocks), dim3(256), bytes, st, t, cw);
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
