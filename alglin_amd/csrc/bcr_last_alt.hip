// Last-block factorizations other than LU (BABD_BorderedCR.cc:1281-1374 with
// select_last LUPQ / QR / QRP), FP64, one workgroup per last block.
//
//   lupq_last_kernel  LU with complete pivoting: the pivot of step k is the
//                     largest |a_ij| of the trailing block, first in
//                     column-major order; rows k <-> P[k], columns k <-> Q[k];
//                     an exactly zero pivot sets status bit 2
//   qr_last_kernel    Householder QR in dgeqr2 order (dlarfg reflectors,
//                     beta = -sign(alpha) * |(alpha, x)|, tau = (beta-alpha)/beta,
//                     v = x / (alpha - beta)), m = Nr >= n = Nc; with PIVOT the
//                     column choice and norm downdating of dlaqp2 (first
//                     largest partial norm, recompute when the downdated norm
//                     loses more than half the digits: tol3z = sqrt(eps)).
//                     jpvt is 0-based.  A zero diagonal of R sets status bit 2,
//                     as a failed factorization of the reference's dense class.
//
// Storage: the block is factored in LDS when Nr*Nc*8 bytes fit, else in place
// in HBM (it stays L2-resident); taus, norms and the column products of a
// reflector live in static LDS arrays of kLastMax entries.
#include "bcr_launch.hh"

namespace bcr {

namespace {

constexpr int kLastMax = 256;   // max Nc handled by the static LDS tables

// block-wide first-max search: value v at linear key `key` (smaller key wins
// a tie); returns the winning key to every thread
__device__ __forceinline__ int block_argmax(double v, int key, double* s_val, int* s_key) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    const double ov = __shfl_xor(v, off);
    const int ok = __shfl_xor(key, off);
    if (ov > v || (ov == v && ok < key)) { v = ov; key = ok; }
  }
  if (lane == 0) { s_val[wid] = v; s_key[wid] = key; }
  __syncthreads();
  if (wid == 0) {
    v = lane < nw ? s_val[lane] : -1.0;
    key = lane < nw ? s_key[lane] : 0x7fffffff;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      const double ov = __shfl_xor(v, off);
      const int ok = __shfl_xor(key, off);
      if (ov > v || (ov == v && ok < key)) { v = ov; key = ok; }
    }
    if (lane == 0) s_key[0] = key;
  }
  __syncthreads();
  const int r = s_key[0];
  __syncthreads();
  return r;
}

template <bool IN_LDS>
__global__ void __launch_bounds__(1024) lupq_last_kernel(double* __restrict__ Hm, int n, int32_t* __restrict__ P,
                                                          int32_t* __restrict__ Q, int32_t* __restrict__ status) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ double s_val[16];
  __shared__ int s_key[16];
  const int tid = threadIdx.x, nth = blockDim.x;
  double* A = IN_LDS ? reinterpret_cast<double*>(smem_raw) : Hm;
  if (IN_LDS)
    for (int i = tid; i < n * n; i += nth) A[i] = Hm[i];
  __syncthreads();
  for (int k = 0; k < n; ++k) {
    const int m = n - k;
    double best = -1.0;
    int bkey = 0;   // valid also when every entry is NaN
    for (int t = tid; t < m * m; t += nth) {   // column-major order over the trailing block
      const double v = fabs(A[(k + t % m) + size_t(k + t / m) * n]);
      if (v > best) { best = v; bkey = t; }
    }
    const int key = block_argmax(best, bkey, s_val, s_key);
    const int pi = k + key % m, pj = k + key / m;
    if (tid == 0) { P[k] = pi; Q[k] = pj; }
    if (pi != k)
      for (int c = tid; c < n; c += nth) {
        const double t = A[k + size_t(c) * n];
        A[k + size_t(c) * n] = A[pi + size_t(c) * n];
        A[pi + size_t(c) * n] = t;
      }
    __syncthreads();
    if (pj != k)
      for (int r = tid; r < n; r += nth) {
        const double t = A[r + size_t(k) * n];
        A[r + size_t(k) * n] = A[r + size_t(pj) * n];
        A[r + size_t(pj) * n] = t;
      }
    __syncthreads();
    const double piv = A[k + size_t(k) * n];
    if (piv == 0.0) {
      if (tid == 0) atomicOr(status, 2);
      break;
    }
    const double rinv = 1.0 / piv;
    for (int r = k + 1 + tid; r < n; r += nth) A[r + size_t(k) * n] *= rinv;
    __syncthreads();
    const int mm = m - 1;
    for (int i = tid; i < mm * mm; i += nth) {
      const int r = k + 1 + i % mm, c = k + 1 + i / mm;
      A[r + size_t(c) * n] -= A[r + size_t(k) * n] * A[k + size_t(c) * n];
    }
    __syncthreads();
  }
  if (IN_LDS)
    for (int i = tid; i < n * n; i += nth) Hm[i] = A[i];
}

template <bool IN_LDS, bool PIVOT>
__global__ void __launch_bounds__(1024) qr_last_kernel(double* __restrict__ Hm, int m, int n, double* __restrict__ tau,
                                                        int32_t* __restrict__ jpvt, int32_t* __restrict__ status) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ double s_val[16];
  __shared__ int s_key[16];
  __shared__ double vn1[kLastMax], vn2[kLastMax], w[kLastMax];
  __shared__ double s_tau, s_beta, s_scal;
  const int tid = threadIdx.x, nth = blockDim.x;
  const int lane = tid & 63, wid = tid >> 6;
  double* A = IN_LDS ? reinterpret_cast<double*>(smem_raw) : Hm;
  auto at = [&](int r, int c) -> double& { return A[r + size_t(c) * m]; };
  if (IN_LDS)
    for (int i = tid; i < m * n; i += nth) A[i] = Hm[i];
  __syncthreads();
  if (PIVOT) {
    for (int c = tid; c < n; c += nth) {
      double s = 0;
      for (int r = 0; r < m; ++r) s += at(r, c) * at(r, c);
      vn1[c] = vn2[c] = sqrt(s);
      jpvt[c] = c;
    }
    __syncthreads();
  }
  const double tol3z = sqrt(2.220446049250313e-16);
  for (int k = 0; k < n; ++k) {
    if (PIVOT) {
      const double v = tid < n - k ? vn1[k + tid] : -1.0;
      const int p = k + block_argmax(v, tid < n - k ? tid : 0x7fffffff, s_val, s_key);
      if (p != k) {
        for (int r = tid; r < m; r += nth) {
          const double t = at(r, k);
          at(r, k) = at(r, p);
          at(r, p) = t;
        }
        if (tid == 0) {
          const int t = jpvt[p]; jpvt[p] = jpvt[k]; jpvt[k] = t;
          vn1[p] = vn1[k];
          vn2[p] = vn2[k];
        }
      }
      __syncthreads();
    }
    // reflector of column k, rows k..m-1
    if (wid == 0) {
      double s = 0;
      for (int r = k + 1 + lane; r < m; r += 64) s += at(r, k) * at(r, k);
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off);
      if (lane == 0) {
        const double alpha = at(k, k), xnorm = sqrt(s);
        if (xnorm == 0.0) {
          s_tau = 0.0; s_beta = alpha; s_scal = 1.0;
        } else {
          const double beta = -copysign(hypot(alpha, xnorm), alpha);
          s_tau = (beta - alpha) / beta;
          s_scal = 1.0 / (alpha - beta);
          s_beta = beta;
        }
        tau[k] = s_tau;
      }
    }
    __syncthreads();
    const double tk = s_tau;
    if (tk != 0.0) {
      for (int r = k + 1 + tid; r < m; r += nth) at(r, k) *= s_scal;
      __syncthreads();
      if (tid == 0) at(k, k) = 1.0;   // v_k = 1 while the reflector is applied
      __syncthreads();
      // w_c = v' A(k:m, c) for the trailing columns, four lanes per column
      const int nc = n - k - 1;
      for (int t = tid; t < 4 * nc + (4 - (4 * nc) % 4) % 4; t += nth) {
        const int c = k + 1 + t / 4, part = t % 4;
        double s = 0;
        if (c < n)
          for (int r = k + part; r < m; r += 4) s += at(r, k) * at(r, c);
        s += __shfl_xor(s, 1);
        s += __shfl_xor(s, 2);
        if (part == 0 && c < n) w[c] = s;
      }
      __syncthreads();
      const int mr = m - k;
      for (int i = tid; i < mr * nc; i += nth) {
        const int r = k + i % mr, c = k + 1 + i / mr;
        at(r, c) -= tk * at(r, k) * w[c];
      }
      __syncthreads();
      if (tid == 0) at(k, k) = s_beta;
      __syncthreads();
    }
    if (tid == 0 && at(k, k) == 0.0) atomicOr(status, 2);
    if (PIVOT) {
      // dlaqp2 partial column norm update
      for (int c = k + 1 + tid; c < n; c += nth) {
        if (vn1[c] != 0.0) {
          double t = fabs(at(k, c)) / vn1[c];
          t = 1.0 - t * t;
          t = t < 0.0 ? 0.0 : t;
          const double r = vn1[c] / vn2[c];
          const double t2 = t * r * r;
          if (t2 <= tol3z) {
            double s = 0;
            for (int rr = k + 1; rr < m; ++rr) s += at(rr, c) * at(rr, c);
            vn1[c] = sqrt(s);
            vn2[c] = vn1[c];
          } else {
            vn1[c] *= sqrt(t);
          }
        }
      }
      __syncthreads();
    }
  }
  if (IN_LDS)
    for (int i = tid; i < m * n; i += nth) Hm[i] = A[i];
}

}  // namespace

hipError_t launch_lupq_last(double* Hm, int Nc, int32_t* P, int32_t* Q, int32_t* status, size_t lds_limit,
                            hipStream_t st) {
  const size_t bytes = size_t(Nc) * Nc * sizeof(double);
  if (bytes <= lds_limit)
    hipLaunchKernelGGL((lupq_last_kernel<true>), dim3(1), dim3(1024), bytes, st, Hm, Nc, P, Q, status);
  else
    hipLaunchKernelGGL((lupq_last_kernel<false>), dim3(1), dim3(1024), 0, st, Hm, Nc, P, Q, status);
  return hipGetLastError();
}

hipError_t launch_qr_last(bool pivot, double* Hm, int Nr, int Nc, double* tau, int32_t* jpvt, int32_t* status,
                          size_t lds_limit, hipStream_t st) {
  if (Nc > kLastMax || Nr < Nc) return hipErrorInvalidValue;
  // static LDS tables: 3 * kLastMax doubles + reduction scratch
  const size_t fixed = 3 * kLastMax * sizeof(double) + 1024;
  const size_t bytes = size_t(Nr) * Nc * sizeof(double);
  const bool in_lds = bytes + fixed <= lds_limit;
  if (pivot) {
    if (in_lds) hipLaunchKernelGGL((qr_last_kernel<true, true>), dim3(1), dim3(1024), bytes, st, Hm, Nr, Nc, tau, jpvt, status);
    else hipLaunchKernelGGL((qr_last_kernel<false, true>), dim3(1), dim3(1024), 0, st, Hm, Nr, Nc, tau, jpvt, status);
  } else {
    if (in_lds) hipLaunchKernelGGL((qr_last_kernel<true, false>), dim3(1), dim3(1024), bytes, st, Hm, Nr, Nc, tau, jpvt, status);
    else hipLaunchKernelGGL((qr_last_kernel<false, false>), dim3(1), dim3(1024), 0, st, Hm, Nr, Nc, tau, jpvt, status);
  }
  return hipGetLastError();
}

}  // namespace bcr
