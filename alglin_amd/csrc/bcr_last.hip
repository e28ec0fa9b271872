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

namespace bcr {


template <typename TW, typename TA>
__global__ void __launch_bounds__(256) assemble_last_kernel(AssembleArgs<TW, TA> a) {
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

// LU with partial pivoting (dgetf2 order: first max, whole-row swaps,
// reciprocal scaling) of the square Nc x Nc last block.
template <bool IN_LDS>
__global__ void __launch_bounds__(1024) lu_last_kernel(double* __restrict__ Hm, int Nc, int32_t* __restrict__ ipiv,
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
      int bi = 0x7fffffff;
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

template <typename TX>
__global__ void __launch_bounds__(256) last_solve_kernel(const double* __restrict__ Hm,
                                                          const int32_t* __restrict__ ipiv, int N, int n, int qr,
                                                          int qx, int nr, int nx, TX* __restrict__ x, int64_t ldx,
                                                          int nrhs) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int Nc = 2 * n + qx + nx;  // == Nr
  const int tid = threadIdx.x, nth = blockDim.x;
  double* A = reinterpret_cast<double*>(smem_raw);
  double* X = A + size_t(Nc) * Nc;
  constexpr int QC = 16;
  for (int i = tid; i < Nc * Nc; i += nth) A[i] = Hm[i];
  const size_t xN = size_t(N) * n;
  const size_t xC = size_t(N + 1) * n + qr;
  for (int q0 = 0; q0 < nrhs; q0 += QC) {
    const int nq = min(QC, nrhs - q0);
    for (int i = tid; i < Nc * nq; i += nth) {
      const int r = i % Nc, q = i / Nc;
      size_t row;
      if (r < n) row = r;
      else if (r < 2 * n + qr) row = xN + (r - n);
      else row = xC + (r - 2 * n - qr);
      X[r + q * Nc] = double(x[row + size_t(q0 + q) * ldx]);
    }
    __syncthreads();
    for (int q = tid; q < nq; q += nth)
      for (int c = 0; c < Nc; ++c) {
        const int pv = ipiv[c] - 1;
        if (pv != c) { const double t = X[c + q * Nc]; X[c + q * Nc] = X[pv + q * Nc]; X[pv + q * Nc] = t; }
      }
    __syncthreads();
    for (int c = 0; c < Nc; ++c) {
      const int m = Nc - c - 1;
      for (int i = tid; i < m * nq; i += nth) {
        const int r = c + 1 + i % m, q = i / m;
        X[r + q * Nc] -= A[r + c * Nc] * X[c + q * Nc];
      }
      __syncthreads();
    }
    for (int c = Nc - 1; c >= 0; --c) {
      const double uinv = 1.0 / A[c + c * Nc];
      for (int i = tid; i < (c + 1) * nq; i += nth) {
        const int r = i % (c + 1), q = i / (c + 1);
        if (r < c) X[r + q * Nc] -= A[r + c * Nc] * (X[c + q * Nc] * uinv);
      }
      __syncthreads();
      for (int q = tid; q < nq; q += nth) X[c + q * Nc] *= uinv;
      __syncthreads();
    }
    for (int i = tid; i < Nc * nq; i += nth) {
      const int r = i % Nc, q = i / Nc;
      size_t row;
      if (r < n) row = r;
      else if (r < 2 * n) row = xN + (r - n);
      else row = size_t(N + 1) * n + (r - 2 * n);
      x[row + size_t(q0 + q) * ldx] = TX(X[r + q * Nc]);
    }
    __syncthreads();
  }
}

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
hipError_t launch_last_solve(const double* Hm, const int32_t* ipiv, int N, int n, int qr, int qx, int nr, int nx,
                             TX* x, int64_t ldx, int nrhs, size_t lds_limit, hipStream_t st) {
  const int Nc = 2 * n + qx + nx;
  const size_t bytes = (size_t(Nc) * Nc + size_t(Nc) * 16) * sizeof(double);
  if (bytes > lds_limit) return hipErrorInvalidConfiguration;
  hipLaunchKernelGGL((last_solve_kernel<TX>), dim3(1), dim3(256), bytes, st, Hm, ipiv, N, n, qr, qx, nr, nx, x, ldx,
                     nrhs);
  return hipGetLastError();
}

template hipError_t launch_assemble_last<double, double>(AssembleArgs<double, double> const&, hipStream_t);
template hipError_t launch_assemble_last<float, float>(AssembleArgs<float, float> const&, hipStream_t);
template hipError_t launch_assemble_last<float, double>(AssembleArgs<float, double> const&, hipStream_t);
template hipError_t launch_last_solve<double>(const double*, const int32_t*, int, int, int, int, int, int, double*,
                                              int64_t, int, size_t, hipStream_t);
template hipError_t launch_last_solve<float>(const double*, const int32_t*, int, int, int, int, int, int, float*,
                                             int64_t, int, size_t, hipStream_t);

}  // namespace bcr
