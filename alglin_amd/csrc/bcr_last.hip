// The bordered last block (BABD_BorderedCR.cc:1281-1479), FP64.
//
// After the cyclic reduction only block row 0 (coupling x_0 and x_N) is left;
// with the boundary rows H and the border rows it forms the Nr x Nc matrix
//
//      n   n  qx  nx
//   [ D0 | E0 | 0 | B0 ]  n        (reduced row 0)
//   [ H0 | HN | Hq| Hp ]  n+qr     (pristine boundary rows)
//   [ C0 | CN | Cq| F  ]  nr       (reduced border rows)
//
// assemble_last_kernel builds it (folding the pending C_N contributions and
// the summed F partials), lu_last_kernel factors it with partial pivoting in
// one workgroup (LDS-resident when it fits, else in place in HBM/L2),
// last_solve_kernel gathers the last-block right-hand side straight from the
// solution vector, solves, and scatters (x_0, x_N, q, xb) back (the
// reference's swap of x[0:n] with x[(N-1)n:Nn], cc:1392-1423, is the same
// gather/scatter written in place).
#include "bcr_launch.hh"
#include "bcr_wave.hh"

namespace bcr {


template <typename TW, typename TA>
__device__ __forceinline__ void assemble_last_body(AssembleArgs<TW, TA> const& a) {
  const int n = a.n, Nr = a.Nr, Nc = a.Nc, qr = a.qr, qx = a.qx, nr = a.nr, nx = a.nx;
  const int tid = threadIdx.x, nth = blockDim.x;
  const size_t rn = size_t(nr) * n;
  for (int i = tid; i < Nr * Nc; i += nth) {
    const int r = i % Nr, c = i / Nr;
    double v = 0;
    if (r < n) {
      if (c < n) v = double(a.D0[r + c * n]);
      else if (c < 2 * n) v = double(a.E0[r + (c - n) * n]);
      else if (c >= 2 * n + qx) v = double(a.B0[r + (c - 2 * n - qx) * n]);
    } else if (r < 2 * n + qr) {
      v = double(a.H[(r - n) + size_t(c) * (n + qr)]);
    } else {
      const int rr = r - 2 * n - qr;
      if (c < n) {
        v = double(a.C[rr + c * nr]);
      } else if (c < 2 * n) {
        const size_t off = size_t(a.N) * rn + rr + (c - n) * nr;
        v = double(a.C[off]) + double(a.Cpend[off]) + double(a.Cpend[size_t(a.N + 1) * rn + off]);
      } else if (c < 2 * n + qx) {
        v = double(a.Cq[rr + (c - 2 * n) * nr]);
      } else {
        v = a.Feff[rr + (c - 2 * n - qx) * nr];
      }
    }
    a.Hm[i] = v;
  }
}

template <typename TW, typename TA>
__global__ void __launch_bounds__(256) assemble_last_kernel(AssembleArgs<TW, TA> a) {
  assemble_last_body<TW, TA>(a);
}

// LU with partial pivoting (dgetf2 order: first max, whole-row swaps,
// reciprocal scaling) of the square Nc x Nc last block.
template <bool IN_LDS>
__device__ __forceinline__ void lu_last_body(double* __restrict__ Hm, int Nc, int32_t* __restrict__ ipiv,
                                             int32_t* __restrict__ status) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ int s_piv;
  const int tid = threadIdx.x, nth = blockDim.x;
  const int lane = tid & 63, wid = tid >> 6;
  const int ld = Nc;
  double* A = IN_LDS ? reinterpret_cast<double*>(smem_raw) : Hm;
  if (IN_LDS)
    for (int i = tid; i < Nc * Nc; i += nth) A[i] = Hm[i];
  __syncthreads();
  for (int c = 0; c < Nc; ++c) {
    if (wid == 0) {
      double best = -1;
      int bi = c;   // a NaN column keeps the diagonal: the index stays valid
      for (int r = c + lane; r < Nc; r += 64) {
        const double v = fabs(A[r + c * ld]);
        if (v > best) { best = v; bi = r; }
      }
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) {
        const double ov = __shfl_xor(best, off);
        const int oi = __shfl_xor(bi, off);
        if (ov > best || (ov == best && oi < bi)) { best = ov; bi = oi; }
      }
      if (lane == 0) {
        s_piv = bi;
        ipiv[c] = bi + 1;
        if (best == 0.0) atomicOr(status, 2);
      }
    }
    __syncthreads();
    const int pv = s_piv;
    if (pv != c)
      for (int s = tid; s < Nc; s += nth) {
        const double t = A[c + s * ld];
        A[c + s * ld] = A[pv + s * ld];
        A[pv + s * ld] = t;
      }
    __syncthreads();
    const double piv = A[c + c * ld];
    const double rinv = piv != 0.0 ? 1.0 / piv : 0.0;
    for (int r = c + 1 + tid; r < Nc; r += nth) A[r + c * ld] *= rinv;
    __syncthreads();
    const int m = Nc - c - 1;
    for (int i = tid; i < m * m; i += nth) {
      const int r = c + 1 + i % m, s = c + 1 + i / m;
      A[r + s * ld] -= A[r + c * ld] * A[c + s * ld];
    }
    __syncthreads();
  }
  if (IN_LDS)
    for (int i = tid; i < Nc * Nc; i += nth) Hm[i] = A[i];
}

// A_LDS: the factored last block is staged in LDS; otherwise (Nr*Nc*8 bytes
// above the LDS limit, e.g. 152 x 152) it is read from global memory, where
// it stays L2-resident, and only the right-hand-side chunk lives in LDS.
// The Nr right-hand-side rows are gathered from x, the Nc solution rows are
// scattered back.  Methods (kLastLU .. kLastQRP):
//   LU   ipiv (1-based, row swaps), L then U            (getrs)
//   LUPQ row swaps P, L, U, column swaps Q in reverse
//   QR   y = Q^T b (reflectors in order), R^{-1} y      (ormqr + trsv)
//   QRP  as QR, then x[jpvt[i]] = y[i]
template <bool IN_LDS>
__global__ void __launch_bounds__(1024) lu_last_kernel(double* __restrict__ Hm, int Nc, int32_t* __restrict__ ipiv,
                                                        int32_t* __restrict__ status) {
  lu_last_body<IN_LDS>(Hm, Nc, ipiv, status);
}

template <typename TX, bool A_LDS>
__device__ __forceinline__ void last_solve_body(LastSolveArgs<TX> const& a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int n = a.n, Nr = a.Nr, Nc = a.Nc, N = a.N, nrhs = a.nrhs;
  const int method = (a.method >= kLastSVD || (a.status && (*a.status & 8))) ? int(kLastSVD) : a.method;
  const int tid = threadIdx.x, nth = blockDim.x, lane = tid & 63, wid = tid >> 6;
  const int ld = Nr;   // column-major Nr x Nc
  const int R = Nr > Nc ? Nr : Nc;
  double* sm = reinterpret_cast<double*>(smem_raw);
  const double* A = A_LDS ? sm : a.Hm;
  double* X = A_LDS ? sm + size_t(Nr) * Nc : sm;   // R x 16
  double* Z = X + size_t(R) * 16;                   // Nc x 16 (QRP permutation)
  constexpr int QC = 16;
  if (A_LDS)
    for (int i = tid; i < Nr * Nc; i += nth) sm[i] = a.Hm[i];
  const size_t xN = size_t(N) * n;
  const size_t xC = size_t(N + 1) * n + a.qr;
  const int64_t ldx = a.ldx;
  for (int q0 = 0; q0 < nrhs; q0 += QC) {
    const int nq = min(QC, nrhs - q0);
    for (int i = tid; i < R * nq; i += nth) {
      const int r = i % R, q = i / R;
      double v = 0;
      if (r < Nr) {
        size_t row;
        if (r < n) row = r;
        else if (r < 2 * n + a.qr) row = xN + (r - n);
        else row = xC + (r - 2 * n - a.qr);
        v = double(a.x[row + size_t(q0 + q) * ldx]);
      }
      X[r + q * R] = v;
    }
    __syncthreads();
    if (method == kLastSVD) {
      // x = V S^+ U^T b over the rank retained singular triplets
      const int k = Nr < Nc ? Nr : Nc, rank = *a.rank;
      for (int i = tid; i < k * nq; i += nth) {
        const int l = i % k, q = i / k;
        double c = 0;
        if (l < rank) {
          for (int r = 0; r < Nr; ++r) c += a.U[r + size_t(l) * Nr] * X[r + q * R];
          c /= a.S[l];
        }
        Z[l + q * Nc] = c;
      }
      __syncthreads();
      for (int i = tid; i < Nc * nq; i += nth) {
        const int r = i % Nc, q = i / Nc;
        double y = 0;
        for (int l = 0; l < rank; ++l) y += a.V[r + size_t(l) * Nc] * Z[l + q * Nc];
        X[r + q * R] = y;
      }
      __syncthreads();
    } else if (method == kLastLU || method == kLastLUPQ) {
      for (int q = tid; q < nq; q += nth)
        for (int c = 0; c < Nc; ++c) {
          const int pv = method == kLastLU ? a.P[c] - 1 : a.P[c];
          if (pv != c) { const double t = X[c + q * R]; X[c + q * R] = X[pv + q * R]; X[pv + q * R] = t; }
        }
      __syncthreads();
      for (int c = 0; c < Nc; ++c) {
        const int m = Nc - c - 1;
        for (int i = tid; i < m * nq; i += nth) {
          const int r = c + 1 + i % m, q = i / m;
          X[r + q * R] -= A[r + c * ld] * X[c + q * R];
        }
        __syncthreads();
      }
    } else {
      // y = Q^T b: H_0 first; wave q forms v' y for column q
      for (int k = 0; k < Nc; ++k) {
        const double tk = a.tau[k];
        if (tk != 0.0) {
          __shared__ double s_w[QC];
          if (wid < nq) {
            double s = 0;
            for (int r = k + lane; r < Nr; r += 64) s += (r == k ? 1.0 : A[r + k * ld]) * X[r + wid * R];
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off);
            if (lane == 0) s_w[wid] = s * tk;
          }
          __syncthreads();
          const int m = Nr - k;
          for (int i = tid; i < m * nq; i += nth) {
            const int r = k + i % m, q = i / m;
            X[r + q * R] -= (r == k ? 1.0 : A[r + k * ld]) * s_w[q];
          }
          __syncthreads();
        }
      }
    }
    // R^{-1} (or U^{-1}), bottom up
    for (int c = method == kLastSVD ? -1 : Nc - 1; c >= 0; --c) {
      const double uinv = 1.0 / A[c + c * ld];
      for (int i = tid; i < c * nq; i += nth) {
        const int r = i % c, q = i / c;
        X[r + q * R] -= A[r + c * ld] * (X[c + q * R] * uinv);
      }
      __syncthreads();
      for (int q = tid; q < nq; q += nth) X[c + q * R] *= uinv;
      __syncthreads();
    }
    const double* Y = X;
    int ldy = R;
    if (method == kLastLUPQ) {
      for (int q = tid; q < nq; q += nth)
        for (int c = Nc - 1; c >= 0; --c) {
          const int pv = a.Q[c];
          if (pv != c) { const double t = X[c + q * R]; X[c + q * R] = X[pv + q * R]; X[pv + q * R] = t; }
        }
      __syncthreads();
    } else if (method == kLastQRP) {
      for (int i = tid; i < Nc * nq; i += nth) {
        const int r = i % Nc, q = i / Nc;
        Z[a.P[r] + q * Nc] = X[r + q * R];
      }
      __syncthreads();
      Y = Z;
      ldy = Nc;
    }
    for (int i = tid; i < Nc * nq; i += nth) {
      const int r = i % Nc, q = i / Nc;
      size_t row;
      if (r < n) row = r;
      else if (r < 2 * n) row = xN + (r - n);
      else row = size_t(N + 1) * n + (r - 2 * n);
      a.x[row + size_t(q0 + q) * ldx] = TX(Y[r + q * ldy]);
    }
    __syncthreads();
  }
}

template <typename TX, bool A_LDS>
__global__ void __launch_bounds__(1024) last_solve_kernel(LastSolveArgs<TX> a) {
  last_solve_body<TX, A_LDS>(a);
}

// ---- batched: one workgroup per system ------------------------------------
template <typename TW, typename TA>
__global__ void __launch_bounds__(256) assemble_last_batched(const AssembleArgs<TW, TA>* args) {
  assemble_last_body<TW, TA>(args[blockIdx.x]);
}

template <bool IN_LDS>
__global__ void __launch_bounds__(1024) lu_last_batched(const LastSolveArgs<double>* args, int32_t* const* status) {
  const LastSolveArgs<double>& a = args[blockIdx.x];
  lu_last_body<IN_LDS>(const_cast<double*>(a.Hm), a.Nc, const_cast<int32_t*>(a.P), status[blockIdx.x]);
}

template <typename TX, bool A_LDS>
__global__ void __launch_bounds__(1024) last_solve_batched(const LastSolveArgs<TX>* args) {
  if (skipped(args[blockIdx.x].skip)) return;
  last_solve_body<TX, A_LDS>(args[blockIdx.x]);
}

// The LU solve of a square last block for one right-hand side, one
// wavefront per system: lane l holds rows l + 64 k (k < RPL) of y, the row
// interchanges are composed into a source row per lane (no serial swaps),
// and the two substitutions walk the columns of the factors from HBM/L2
// eight at a time with y[c] broadcast by v_readlane, without a workgroup
// barrier.  Same arithmetic per element as last_solve_body's LU path.
template <typename TX, int RPL>
__global__ void __launch_bounds__(64) last_solve_lu_wave_batched(const LastSolveArgs<TX>* args) {
  const LastSolveArgs<TX>& a = args[blockIdx.x];
  if (skipped(a.skip)) return;
  constexpr int CH = 8;
  const int lane = threadIdx.x;
  const int n = a.n, Nc = a.Nc, ld = a.Nr;
  const double* A = a.Hm;
  const size_t xN = size_t(a.N) * n, xC = size_t(a.N + 1) * n + a.qr;
  TX* const x = a.x;
  // the source row of each row after the interchanges c = 0 .. Nc-1
  // (1-based ipiv), found by undoing them from the last
  int pos[RPL];
#pragma unroll
  for (int k = 0; k < RPL; ++k) pos[k] = lane + 64 * k;
  for (int c = Nc - 1; c >= 0; --c) {
    const int pv = a.P[c] - 1;
#pragma unroll
    for (int k = 0; k < RPL; ++k) pos[k] = pos[k] == c ? pv : pos[k] == pv ? c : pos[k];
  }
  double y[RPL], dinv[RPL];
#pragma unroll
  for (int k = 0; k < RPL; ++k) {
    const int r = lane + 64 * k, s = r < Nc ? pos[k] : 0;
    const size_t row = s < n ? size_t(s) : s < 2 * n + a.qr ? xN + (s - n) : xC + (s - 2 * n - a.qr);
    y[k] = r < Nc ? double(x[row]) : 0.0;
    dinv[k] = r < Nc ? 1.0 / A[r + size_t(r) * ld] : 1.0;
  }
  // y <- L^{-1} y (unit lower), column c of L below the diagonal
#pragma unroll
  for (int kc = 0; kc < RPL; ++kc) {
    const int cend = min(64 * kc + 64, Nc);
    for (int c0 = 64 * kc; c0 < cend; c0 += CH) {
      double l[CH][RPL];
#pragma unroll
      for (int i = 0; i < CH; ++i)
#pragma unroll
        for (int k = 0; k < RPL; ++k) {
          const int c = min(c0 + i, Nc - 1), r = lane + 64 * k;
          l[i][k] = A[(r < Nc ? r : 0) + size_t(c) * ld];
        }
#pragma unroll
      for (int i = 0; i < CH; ++i) {
        const int c = c0 + i;
        if (c < cend) {
          const double yc = wave::rl(y[kc], c - 64 * kc);
#pragma unroll
          for (int k = 0; k < RPL; ++k) {
            const int r = lane + 64 * k;
            if (r > c && r < Nc) y[k] -= l[i][k] * yc;
          }
        }
      }
    }
  }
  // y <- U^{-1} y, bottom up: y_c = y_c / u_cc, then the rows above
#pragma unroll
  for (int kc = RPL - 1; kc >= 0; --kc) {
    const int cbeg = 64 * kc;
    for (int c1 = min(64 * kc + 63, Nc - 1); c1 >= cbeg; c1 -= CH) {
      double u[CH][RPL];
#pragma unroll
      for (int i = 0; i < CH; ++i)
#pragma unroll
        for (int k = 0; k < RPL; ++k) {
          const int c = max(c1 - i, 0), r = lane + 64 * k;
          u[i][k] = A[(r < Nc ? r : 0) + size_t(c) * ld];
        }
#pragma unroll
      for (int i = 0; i < CH; ++i) {
        const int c = c1 - i;
        if (c >= cbeg) {
          const double xc = wave::rl(y[kc] * dinv[kc], c - 64 * kc);
#pragma unroll
          for (int k = 0; k < RPL; ++k) {
            const int r = lane + 64 * k;
            y[k] = r < c ? y[k] - u[i][k] * xc : r == c ? xc : y[k];
          }
        }
      }
    }
  }
#pragma unroll
  for (int k = 0; k < RPL; ++k) {
    const int r = lane + 64 * k;
    if (r < Nc) {
      const size_t row = r < n ? size_t(r) : r < 2 * n ? xN + (r - n) : size_t(a.N + 1) * n + (r - 2 * n);
      x[row] = TX(y[k]);
    }
  }
}

template <typename TW, typename TA>
hipError_t launch_assemble_last_batched(const AssembleArgs<TW, TA>* args, int nsys, hipStream_t st) {
  hipLaunchKernelGGL((assemble_last_batched<TW, TA>), dim3(nsys), dim3(256), 0, st, args);
  return hipGetLastError();
}

// The same factorization (Nc <= 64 RPL) by one workgroup of NWV wavefronts
// per system with the block in registers: wavefront w holds columns
// [CW w, CW w + CW) of every row (lane l: rows l + 64 k).  Rows stay in
// their lanes with their current positions tracked (the pivot row takes
// position c, the row there the pivot's old place); the wavefront owning
// column c searches it (largest |a|, ties to the smaller position, as
// lu_last_body's scan) and publishes the multipliers through LDS, and every
// wavefront applies the rank-1 update to its columns after c.  The rows are
// stored back in their final positions with the 1-based ipiv, so the result
// and its arithmetic per element are lu_last_body's.
template <int RPL, int NWV>
__global__ void __launch_bounds__(64 * NWV) lu_last_regs_batched(const LastSolveArgs<double>* args,
                                                                 int32_t* const* status) {
  using V16d = double __attribute__((ext_vector_type(16)));
  constexpr int CW = 64 * RPL / NWV, CB = CW / 16;   // columns and 16-column blocks per wavefront
  static_assert(CW * NWV == 64 * RPL && CW % 16 == 0, "the waves cover the columns in 16-column blocks");
  __shared__ double lbuf[2][64 * RPL];
  __shared__ int pbuf[2][2];
  const LastSolveArgs<double>& a = args[blockIdx.x];
  const int Nc = a.Nc;
  double* const Hm = const_cast<double*>(a.Hm);
  int32_t* const ipiv = const_cast<int32_t*>(a.P);
  int32_t* const st = status[blockIdx.x];
  const int lane = threadIdx.x & 63;
  const int w = NWV == 1 ? 0 : __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
  V16d V[RPL][CB];
  int pos[RPL];
#pragma unroll
  for (int k = 0; k < RPL; ++k) {
    const int r = lane + 64 * k;
    pos[k] = r;
#pragma unroll
    for (int jj = 0; jj < CW; ++jj) {
      const int gj = CW * w + jj;
      V[k][jj / 16][jj % 16] = r < Nc && gj < Nc ? Hm[r + size_t(gj) * Nc] : 0.0;
    }
  }
  // column blocks of 16 (unrolled: the owning wavefront and its block are
  // compile-time per block), steps inside a block rolled (the column read
  // through a uniform register index)
#pragma unroll
  for (int gb = 0; gb < 4 * RPL; ++gb) {
    const int o = 16 * gb / CW, lb = (16 * gb % CW) / 16;
    const int cend = min(16 * gb + 16, Nc);
    for (int c = 16 * gb; c < cend; ++c) {
      const int cc = c - 16 * gb;
      double lk[RPL];
      int who = 0, ps = 0, P = 0;
      if (w == o) {
        double x[RPL];
        bool cb = false;
        unsigned hb = 0, lbits = 0;
        int pb = 0, kb = 0;
#pragma unroll
        for (int k = 0; k < RPL; ++k) {
          x[k] = V[k][lb][cc];
          unsigned h, l;
          wave::abs_key(x[k], h, l);
          if (x[k] != x[k]) h = l = 0;   // a NaN is never preferred
          const bool cand = pos[k] >= c && lane + 64 * k < Nc;
          const bool better = cand && (!cb || h > hb || (h == hb && (l > lbits || (l == lbits && pos[k] < pb))));
          if (better) {
            cb = true;
            hb = h;
            lbits = l;
            pb = pos[k];
            kb = k;
          }
        }
        bool zero;
        who = wave::argmax_pos<true>(cb, hb, lbits, pb, lane, zero);
        P = __builtin_amdgcn_readlane(pb, who);
        ps = __builtin_amdgcn_readlane(kb, who);
        double xs = x[0];
#pragma unroll
        for (int k = 1; k < RPL; ++k) xs = kb == k ? x[k] : xs;
        const double pv = wave::rl(xs, who);
        if (lane == 0) {
          if (zero) atomicOr(st, 2);
          ipiv[c] = P + 1;
        }
        const double rinv = pv != 0.0 ? 1.0 / pv : 0.0;
#pragma unroll
        for (int k = 0; k < RPL; ++k) {
          const int np = pos[k] == c ? P : pos[k] == P ? c : pos[k];
          const bool below = np > c && lane + 64 * k < Nc;
          const double m = x[k] * rinv;
          lk[k] = below ? m : 0.0;
          V[k][lb][cc] = below ? m : x[k];
          if constexpr (NWV > 1) lbuf[c & 1][lane + 64 * k] = lk[k];
        }
        if (NWV > 1 && lane == 0) {
          pbuf[c & 1][0] = who | (ps << 8);
          pbuf[c & 1][1] = P;
        }
      }
      if constexpr (NWV > 1) {
        __syncthreads();
        if (w != o) {
          const int q = __builtin_amdgcn_readfirstlane(pbuf[c & 1][0]);
          who = q & 255;
          ps = q >> 8;
          P = __builtin_amdgcn_readfirstlane(pbuf[c & 1][1]);
#pragma unroll
          for (int k = 0; k < RPL; ++k) lk[k] = lbuf[c & 1][lane + 64 * k];
        }
      }
#pragma unroll
      for (int k = 0; k < RPL; ++k) pos[k] = pos[k] == c ? P : pos[k] == P ? c : pos[k];
      // this wavefront's columns after c: A[r][j] -= l_r * A[c][j]
      if (CW * w + CW - 1 > c) {
#pragma unroll
        for (int jj = 0; jj < CW; ++jj) {
          if (CW * w + jj > c) {
            double v = V[0][jj / 16][jj % 16];
#pragma unroll
            for (int k = 1; k < RPL; ++k) v = ps == k ? V[k][jj / 16][jj % 16] : v;
            const double u = wave::rl(v, who);
#pragma unroll
            for (int k = 0; k < RPL; ++k) V[k][jj / 16][jj % 16] -= lk[k] * u;
          }
        }
      }
    }
  }
#pragma unroll
  for (int k = 0; k < RPL; ++k) {
    if (lane + 64 * k < Nc) {
#pragma unroll
      for (int jj = 0; jj < CW; ++jj) {
        const int gj = CW * w + jj;
        if (gj < Nc) Hm[pos[k] + size_t(gj) * Nc] = V[k][jj / 16][jj % 16];
      }
    }
  }
}

// args[s].Hm / .P (ipiv) / .Nc describe the square last block of system s;
// Ncmax is the largest of the launch
hipError_t launch_lu_last_batched(const LastSolveArgs<double>* args, int32_t* const* status, int nsys, int Ncmax,
                                  size_t lds_limit, hipStream_t st) {
  if (Ncmax <= 64) {
    hipLaunchKernelGGL((lu_last_regs_batched<1, 1>), dim3(nsys), dim3(64), 0, st, args, status);
    return hipGetLastError();
  }
  if (Ncmax <= 128) {
    hipLaunchKernelGGL((lu_last_regs_batched<2, 4>), dim3(nsys), dim3(256), 0, st, args, status);
    return hipGetLastError();
  }
  const size_t bytes = size_t(Ncmax) * Ncmax * sizeof(double);
  if (bytes <= lds_limit)
    hipLaunchKernelGGL((lu_last_batched<true>), dim3(nsys), dim3(1024), bytes, st, args, status);
  else
    hipLaunchKernelGGL((lu_last_batched<false>), dim3(nsys), dim3(1024), 0, st, args, status);
  return hipGetLastError();
}

template <typename TX>
hipError_t launch_last_solve_batched(const LastSolveArgs<TX>* args, int nsys, int Nrmax, int Ncmax, size_t lds_limit,
                                     hipStream_t st) {
  const int R = Nrmax > Ncmax ? Nrmax : Ncmax;
  const size_t xbytes = (size_t(R) + Ncmax) * 16 * sizeof(double);
  const size_t bytes = size_t(Nrmax) * Ncmax * sizeof(double) + xbytes;
  if (xbytes > lds_limit) return hipErrorInvalidConfiguration;
  if (bytes <= lds_limit)
    hipLaunchKernelGGL((last_solve_batched<TX, true>), dim3(nsys), dim3(1024), bytes, st, args);
  else
    hipLaunchKernelGGL((last_solve_batched<TX, false>), dim3(nsys), dim3(1024), xbytes, st, args);
  return hipGetLastError();
}

template hipError_t launch_assemble_last_batched<float, double>(const AssembleArgs<float, double>*, int, hipStream_t);
template hipError_t launch_assemble_last_batched<double, double>(const AssembleArgs<double, double>*, int, hipStream_t);
template hipError_t launch_last_solve_batched<double>(const LastSolveArgs<double>*, int, int, int, size_t, hipStream_t);
// the wavefront LU solve above, for square last blocks with Nc <= 256 solved
// by LU for one right-hand side; false when it does not apply
template <typename TX>
bool launch_last_solve_lu_wave_batched(const LastSolveArgs<TX>* args, int nsys, int Ncmax, hipStream_t st,
                                       hipError_t* err) {
  if (Ncmax < 1 || Ncmax > 256) return false;
  if (Ncmax <= 64) hipLaunchKernelGGL((last_solve_lu_wave_batched<TX, 1>), dim3(nsys), dim3(64), 0, st, args);
  else if (Ncmax <= 128) hipLaunchKernelGGL((last_solve_lu_wave_batched<TX, 2>), dim3(nsys), dim3(64), 0, st, args);
  else hipLaunchKernelGGL((last_solve_lu_wave_batched<TX, 4>), dim3(nsys), dim3(64), 0, st, args);
  *err = hipGetLastError();
  return true;
}
template bool launch_last_solve_lu_wave_batched<double>(const LastSolveArgs<double>*, int, int, hipStream_t,
                                                        hipError_t*);


template <typename TW, typename TA>
hipError_t launch_assemble_last(AssembleArgs<TW, TA> const& a, hipStream_t st) {
  hipLaunchKernelGGL((assemble_last_kernel<TW, TA>), dim3(1), dim3(256), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_lu_last(double* Hm, int Nc, int32_t* ipiv, int32_t* status, size_t lds_limit, hipStream_t st) {
  const size_t bytes = size_t(Nc) * Nc * sizeof(double);
  if (bytes <= lds_limit)
    hipLaunchKernelGGL((lu_last_kernel<true>), dim3(1), dim3(1024), bytes, st, Hm, Nc, ipiv, status);
  else
    hipLaunchKernelGGL((lu_last_kernel<false>), dim3(1), dim3(1024), 0, st, Hm, Nc, ipiv, status);
  return hipGetLastError();
}

template <typename TX>
hipError_t launch_last_solve(LastSolveArgs<TX> const& a, size_t lds_limit, hipStream_t st) {
  const int R = a.Nr > a.Nc ? a.Nr : a.Nc;
  const size_t xbytes = (size_t(R) + a.Nc) * 16 * sizeof(double);
  const size_t bytes = size_t(a.Nr) * a.Nc * sizeof(double) + xbytes;
  if (xbytes > lds_limit) return hipErrorInvalidConfiguration;
  if (bytes <= lds_limit)
    hipLaunchKernelGGL((last_solve_kernel<TX, true>), dim3(1), dim3(1024), bytes, st, a);
  else
    hipLaunchKernelGGL((last_solve_kernel<TX, false>), dim3(1), dim3(1024), xbytes, st, a);
  return hipGetLastError();
}

template hipError_t launch_assemble_last<double, double>(AssembleArgs<double, double> const&, hipStream_t);
template hipError_t launch_assemble_last<float, float>(AssembleArgs<float, float> const&, hipStream_t);
template hipError_t launch_assemble_last<float, double>(AssembleArgs<float, double> const&, hipStream_t);
template hipError_t launch_last_solve<double>(LastSolveArgs<double> const&, size_t, hipStream_t);
template hipError_t launch_last_solve<float>(LastSolveArgs<float> const&, size_t, hipStream_t);

}  // namespace bcr
