// Pair elimination of one cyclic-reduction level: generic workgroup kernel.
//
// One workgroup owns one pair (jp, j, jpp) of level L and holds the augmented
// panel of BABD_BorderedCR.cc:1090-1154 in LDS (or in a per-pair HBM scratch
// when it does not fit), column-major with R = 2n+nr rows and W = 3n+nx cols:
//
//            cols:  x_j (T)   x_jp (D)   x_jpp (E)   xb (B)
//   rows 0..n-1  : [ E_jp   |  D_jp    |   0      | B_jp ]   row block jp
//   rows n..2n-1 : [ D_j    |  0       |   E_j    | B_j  ]   row block j
//   rows 2n..    : [ C_j    |  C_jp    |   0      |  0   ]   border rows
//
// n elimination steps on the first n columns (LU with partial pivoting among
// the first 2n rows; Householder QR; QR with column pivoting) turn it into
//
//   rows 0..n-1  : [ U or R | D_j'  | E_j'  | B_j' ]  -> pivot row j (cc:979-980)
//   rows n..2n-1 : [ L or V | D_jp' | E_jp' | B_jp'] -> reduced row jp
//   rows 2n..    : [ C_j U^-1 | C_jp - C_j' D_j' | -C_j' E_j' | -C_j' B_j' ]
//
// which is exactly buildT + applyT x3 + the C/F updates of the reference:
// getrf on [E_jp;D_j] followed by laswp/trsm/gemm on the other columns is the
// same sequence of row operations as right-looking elimination of the whole
// augmented row, and the border rows are eliminated with the final pivot rows
// (trsm 'R','U' + two gemm, cc:1116-1143) without taking part in pivoting.
//
// The -C_j' E_j' block is the contribution to C_jpp and goes to the pending
// slot Cpend[slot_cur][jpp]; -C_j' B_j' goes to Fpart[j].
//
// Pending slots (level stride L, slot s(L) = log2(L)&1).  All slots are
// cleared before the first level; writers accumulate (+=).
//   - the only writer of Cpend[s(L)][x] at level L is the pair with jpp == x;
//   - the pair with j == x reads both slots of x.  x is an odd multiple of L:
//     it is never a jpp at L (no concurrent writer) and, once eliminated,
//     never a jpp nor read again, so its slots need no clearing;
//   - the pair with jp == x reads Cpend[s(L/2)][x] (written at L/2) and
//     clears it, because that slot is written again two levels later.  A
//     block left unpaired at L/2 satisfies x + L/2 >= N and so can never be a
//     jp at L: nothing older than L/2 is pending for a jp, and an unpaired
//     block's slot keeps accumulating until the block is read as a j;
//   - x = N is never read by a pair; the last-block assembly reads it.
#include "bcr_launch.hh"

namespace bcr {

// wave-level argmax over (value, index); ties keep the smaller index
template <typename T>
__device__ __forceinline__ void wave_argmax(T& v, int& idx) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    T ov = __shfl_xor(v, off);
    int oi = __shfl_xor(idx, off);
    if (ov > v || (ov == v && oi < idx)) { v = ov; idx = oi; }
  }
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

template <typename TS, typename TW, int METHOD, bool IN_LDS>
__device__ __forceinline__ void pair_generic_body(FactorArgs<TS, TW> a, TW* __restrict__ gscratch, int bid) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ TW s_vn1[128], s_vn2[128], s_tau[64], s_lr[16];
  __shared__ TW s_red[4];
  __shared__ int s_piv;
  __shared__ int s_jpvt[64];

  const int p = bid;
  if (p >= a.npairs) return;
  int j, jp, jpp;
  pair_of_uniform(a, p, j, jp, jpp);
  const int n = a.n, nr = a.nr, nx = a.nx;
  const int n2 = 2 * n;
  const int R = n2 + nr;
  const int Wc = 3 * n + nx;
  const int ld = R;
  const int tid = threadIdx.x, nth = blockDim.x;
  const int lane = tid & (kWave - 1), wid = tid / kWave;
  TW* P = IN_LDS ? reinterpret_cast<TW*>(smem_raw) : gscratch + size_t(p) * ld * Wc;

  const size_t nn = size_t(n) * n;
  const size_t rn = size_t(nr) * n;
  const int sp = a.slot_cur ^ 1;
  TW* pend_cur = a.Cpend + size_t(a.slot_cur) * (a.N + 1) * rn;
  TW* pend_prv = a.Cpend + size_t(sp) * (a.N + 1) * rn;

  // ---- load the augmented panel ------------------------------------------
  for (int i = tid; i < n * n; i += nth) {
    const int r = i % n, c = i / n;
    P[r + c * ld] = TW(a.sE[jp * nn + i]);                  // E_jp  (T, top)
    P[n + r + c * ld] = TW(a.sD[j * nn + i]);               // D_j   (T, bottom)
    P[r + (n + c) * ld] = TW(a.sD[jp * nn + i]);            // D_jp
    P[n + r + (n + c) * ld] = TW(0);
    P[r + (n2 + c) * ld] = TW(0);
    P[n + r + (n2 + c) * ld] = TW(a.sE[j * nn + i]);        // E_j
  }
  for (int i = tid; i < n * nx; i += nth) {
    const int r = i % n, c = i / n;
    P[r + (3 * n + c) * ld] = TW(a.sB[jp * size_t(n) * nx + i]);
    P[n + r + (3 * n + c) * ld] = TW(a.sB[j * size_t(n) * nx + i]);
  }
  for (int i = tid; i < nr * n; i += nth) {
    const int r = i % nr, c = i / nr;
    TW cj = TW(a.sC[j * rn + i]);
    TW cjp = TW(a.sC[jp * rn + i]);
    if (!a.first) {
      cj += pend_prv[j * rn + i] + pend_cur[j * rn + i];
      cjp += pend_prv[jp * rn + i];
      pend_prv[jp * rn + i] = TW(0);
    }
    P[n2 + r + c * ld] = cj;
    P[n2 + r + (n + c) * ld] = cjp;
    P[n2 + r + (n2 + c) * ld] = TW(0);
  }
  for (int i = tid; i < nr * nx; i += nth) {
    const int r = i % nr, c = i / nr;
    P[n2 + r + (3 * n + c) * ld] = TW(0);
  }
  if (METHOD == kQRP) {
    for (int c = tid; c < n; c += nth) {
      TW s = 0;
      for (int r = 0; r < n2; ++r) { TW v = P[r + c * ld]; s += v * v; }
      s_vn1[c] = sqrt(s);
      s_vn2[c] = s_vn1[c];
      s_jpvt[c] = c;
    }
  }
  __syncthreads();

  // ---- n elimination steps --------------------------------------------------
  for (int c = 0; c < n; ++c) {
    if (METHOD == kLU) {
      if (wid == 0) {
        // a NaN never compares greater: the default pivot is the diagonal, so
        // the index stays valid whatever the panel holds
        TW best = TW(-1);
        int bi = c;
        for (int r = c + lane; r < n2; r += kWave) {
          TW v = dev_abs(P[r + c * ld]);
          if (v > best) { best = v; bi = r; }
        }
        wave_argmax(best, bi);
        if (lane == 0) {
          s_piv = bi;
          if (best == TW(0)) atomicOr(a.status, 1);
        }
      }
      __syncthreads();
      const int piv = s_piv;
      if (piv != c)
        for (int s = tid; s < Wc; s += nth) {
          TW t = P[c + s * ld];
          P[c + s * ld] = P[piv + s * ld];
          P[piv + s * ld] = t;
        }
      if (tid == 0) s_jpvt[c] = piv;
      __syncthreads();
      const TW pv = P[c + c * ld];
      const TW rinv = pv != TW(0) ? TW(1) / pv : TW(0);
      for (int r = c + 1 + tid; r < R; r += nth) P[r + c * ld] *= rinv;
      __syncthreads();
      const int rows = R - c - 1, cols = Wc - c - 1;
      for (int i = tid; i < rows * cols; i += nth) {
        const int r = c + 1 + i % rows, s = c + 1 + i / rows;
        P[r + s * ld] -= P[r + c * ld] * P[c + s * ld];
      }
      __syncthreads();
    } else {
      if (METHOD == kQRP) {
        if (wid == 0) {
          TW best = TW(-1);
          int bi = c;   // valid also when every norm is NaN
          for (int s = c + lane; s < n; s += kWave) {
            if (s_vn1[s] > best) { best = s_vn1[s]; bi = s; }
          }
          wave_argmax(best, bi);
          if (lane == 0) s_piv = bi;
        }
        __syncthreads();
        const int pvt = s_piv;
        if (pvt != c) {
          for (int r = tid; r < R; r += nth) {   // whole column incl. border rows
            TW t = P[r + c * ld];
            P[r + c * ld] = P[r + pvt * ld];
            P[r + pvt * ld] = t;
          }
          if (tid == 0) {
            int t = s_jpvt[c]; s_jpvt[c] = s_jpvt[pvt]; s_jpvt[pvt] = t;
            s_vn1[pvt] = s_vn1[c];
            s_vn2[pvt] = s_vn2[c];
          }
        }
        __syncthreads();
      }
      // Householder reflector for column c (dlarfg)
      if (wid == 0) {
        TW ss = 0;
        for (int r = c + 1 + lane; r < n2; r += kWave) { TW v = P[r + c * ld]; ss += v * v; }
        ss = wave_sum(ss);
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
      // apply H = I - tau v v^T to columns c+1 .. Wc-1, rows c .. 2n-1
      {
        const TW tau = s_tau[c];
        if (tau != TW(0))
          for (int s = c + 1 + tid; s < Wc; s += nth) {
            TW w = P[c + s * ld];
            for (int r = c + 1; r < n2; ++r) w += P[r + c * ld] * P[r + s * ld];
            w *= tau;
            P[c + s * ld] -= w;
            for (int r = c + 1; r < n2; ++r) P[r + s * ld] -= P[r + c * ld] * w;
          }
      }
      __syncthreads();
      // row c is final: eliminate it from the border rows, update norms
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
      for (int i = tid; i < nr * (Wc - c - 1); i += nth) {
        const int r = i % nr, s = c + 1 + i / nr;
        P[n2 + r + s * ld] -= s_lr[r] * P[c + s * ld];
      }
      if (METHOD == kQRP) {
        // dlaqp2 norm downdating (LAWN 176)
        const TW tol3z = sqrt(TW(METHOD == kQRP && sizeof(TW) == 8 ? 2.220446049250313e-16 : 1.1920929e-07));
        for (int s = c + 1 + tid; s < n; s += nth) {
          if (s_vn1[s] != TW(0)) {
            TW t = dev_abs(P[c + s * ld]) / s_vn1[s];
            t = TW(1) - t * t;
            t = t < TW(0) ? TW(0) : t;
            const TW q = s_vn1[s] / s_vn2[s];
            if (t * q * q <= tol3z) {
              TW ssum = 0;
              for (int r = c + 1; r < n2; ++r) { TW v = P[r + s * ld]; ssum += v * v; }
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
  }

  // ---- write back -----------------------------------------------------------
  TW* Tj = a.Tm + size_t(j) * tsize(n);
  for (int i = tid; i < n2 * n; i += nth) {
    const int r = i % n2, c = i / n2;
    Tj[i] = P[r + c * ld];
  }
  if (METHOD != kLU)
    for (int c = tid; c < n; c += nth) Tj[size_t(n2) * n + c] = s_tau[c];
  if (tid == 0) {
    int32_t* Pj = a.perm + size_t(j) * n;
    if (METHOD == kLU) {
      for (int c = 0; c < n; ++c) Pj[c] = s_jpvt[c] + 1;
    } else if (METHOD == kQRP) {
      // permutation -> exchange list (BABD_BorderedCR.hxx:265-275)
      int Pm[64];
      for (int c = 0; c < n; ++c) Pm[c] = s_jpvt[c] + 1;
      for (int i = 0; i < n; ++i) {
        int jj = i;
        while (jj < n && Pm[jj] != i + 1) ++jj;
        int t = Pm[jj]; Pm[jj] = Pm[i]; Pm[i] = t;
        Pj[i] = jj;
      }
    }
  }
  for (int i = tid; i < n * n; i += nth) {
    const int r = i % n, c = i / n;
    a.D[j * nn + i] = P[r + (n + c) * ld];
    a.D[jp * nn + i] = P[n + r + (n + c) * ld];
    a.E[j * nn + i] = P[r + (n2 + c) * ld];
    a.E[jp * nn + i] = P[n + r + (n2 + c) * ld];
  }
  for (int i = tid; i < n * nx; i += nth) {
    const int r = i % n, c = i / n;
    a.B[j * size_t(n) * nx + i] = P[r + (3 * n + c) * ld];
    a.B[jp * size_t(n) * nx + i] = P[n + r + (3 * n + c) * ld];
  }
  for (int i = tid; i < nr * n; i += nth) {
    const int r = i % nr, c = i / nr;
    a.C[j * rn + i] = P[n2 + r + c * ld];
    a.C[jp * rn + i] = P[n2 + r + (n + c) * ld];
    pend_cur[jpp * rn + i] += P[n2 + r + (n2 + c) * ld];
  }
  for (int i = tid; i < nr * nx; i += nth) {
    const int r = i % nr, c = i / nr;
    a.Fpart[size_t(j) * nr * nx + i] = P[n2 + r + (3 * n + c) * ld];
  }
}

template <typename TS, typename TW, int METHOD, bool IN_LDS>
__global__ void __launch_bounds__(256) pair_generic_kernel(FactorArgs<TS, TW> a, TW* __restrict__ gscratch) {
  pair_generic_body<TS, TW, METHOD, IN_LDS>(a, gscratch, blockIdx.x);
}


template <typename TS, typename TW>
hipError_t launch_pair_generic(int method, FactorArgs<TS, TW> const& a, TW* gscratch, size_t lds_limit,
                               hipStream_t st) {
  const int R = 2 * a.n + a.nr, Wc = 3 * a.n + a.nx;
  const size_t bytes = size_t(R) * Wc * sizeof(TW);
  dim3 grid(a.npairs), block(256);
  const bool lds = bytes <= lds_limit;
  if (!lds && gscratch == nullptr) return hipErrorInvalidValue;
#define BCR_LAUNCH(M)                                                                                    \
  if (lds) hipLaunchKernelGGL((pair_generic_kernel<TS, TW, M, true>), grid, block, bytes, st, a, gscratch); \
  else     hipLaunchKernelGGL((pair_generic_kernel<TS, TW, M, false>), grid, block, 0, st, a, gscratch);
  if (method == kLU) { BCR_LAUNCH(kLU) }
  else if (method == kQR) { BCR_LAUNCH(kQR) }
  else { BCR_LAUNCH(kQRP) }
#undef BCR_LAUNCH
  return hipGetLastError();
}

template <typename TS, typename TW, int METHOD>
__global__ void __launch_bounds__(256) pair_generic_batched(BatchTable<FactorArgs<TS, TW>> t) {
  const int e = t.entry[blockIdx.x];
  pair_generic_body<TS, TW, METHOD, true>(t.args[e], nullptr, int(blockIdx.x) - t.first[e]);
}

// panels staged in LDS only: `bytes` is the largest panel of the launch
template <typename TS, typename TW>
hipError_t launch_pair_generic_batched(int method, BatchTable<FactorArgs<TS, TW>> const& t, int nblocks, size_t bytes,
                                       size_t lds_limit, hipStream_t st) {
  if (bytes > lds_limit) return hipErrorInvalidValue;
  if (method == kLU) hipLaunchKernelGGL((pair_generic_batched<TS, TW, kLU>), dim3(nblocks), dim3(256), bytes, st, t);
  else if (method == kQR) hipLaunchKernelGGL((pair_generic_batched<TS, TW, kQR>), dim3(nblocks), dim3(256), bytes, st, t);
  else hipLaunchKernelGGL((pair_generic_batched<TS, TW, kQRP>), dim3(nblocks), dim3(256), bytes, st, t);
  return hipGetLastError();
}

template hipError_t launch_pair_generic_batched<double, double>(int, BatchTable<FactorArgs<double, double>> const&, int,
                                                                size_t, size_t, hipStream_t);
template hipError_t launch_pair_generic_batched<float, float>(int, BatchTable<FactorArgs<float, float>> const&, int,
                                                              size_t, size_t, hipStream_t);
template hipError_t launch_pair_generic_batched<double, float>(int, BatchTable<FactorArgs<double, float>> const&, int,
                                                               size_t, size_t, hipStream_t);
template hipError_t launch_pair_generic<double, double>(int, FactorArgs<double, double> const&, double*, size_t, hipStream_t);
template hipError_t launch_pair_generic<float, float>(int, FactorArgs<float, float> const&, float*, size_t, hipStream_t);
template hipError_t launch_pair_generic<double, float>(int, FactorArgs<double, float> const&, float*, size_t, hipStream_t);

}  // namespace bcr
