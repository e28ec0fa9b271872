// ORACLE — test infrastructure only. Nothing in the product links this file.
//
// Dense column-major kernels used by the CPU restatement of BorderedCR.
// The reference delegates these to lapack_wrapper (an un-vendored submodule
// over OpenBLAS + reference LAPACK; called at BABD_BorderedCR.cc:856-893,
// 939-975, 1004-1038, 1116-1153).  Each routine below restates the published
// LAPACK/BLAS algorithm of the routine the reference calls:
//
//   getrf  -> LAPACK dgetf2 (unblocked right-looking LU, idamax = first max,
//             whole-row swaps, multipliers scaled by the reciprocal pivot)
//   laswp  -> dlaswp, forward interchanges, 1-based ipiv
//   geqrf  -> dgeqr2 + dlarfg (beta = -sign(alpha)*||[alpha;x]||,
//             tau = (beta-alpha)/beta, v = x/(alpha-beta))
//   ormqr  -> dorm2r, SIDE='L', TRANS='T' (H_0 applied first)
//   geqp3  -> dlaqp2 with free columns (column-norm pivoting, idamax = first
//             max, LAWN 176 norm downdating with tol3z = sqrt(eps))
//   trsm / gemm / gemv / trsv -> the BLAS reference loop orders
//
// For 2n x n panels with n <= 64 reference LAPACK (which OpenBLAS ships for
// geqrf/geqp3/ormqr) takes exactly these unblocked code paths, because
// min(m,n) is below the blocking crossover NX=128.  Bitwise agreement with
// OpenBLAS's own getrf and BLAS kernels is not claimed: parity is
// tolerance-based, see the tests that pin this file.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

namespace oracle {

using integer = int32_t;

template <class T> inline T sq(T v) { return v * v; }

// first index of max |x[i]| (BLAS i?amax semantics, 0-based result)
template <class T>
inline integer iamax(integer n, T const* x, integer inc) {
  if (n <= 0) return -1;
  integer p = 0;
  T vmax = std::abs(x[0]);
  for (integer i = 1; i < n; ++i) {
    T v = std::abs(x[i * inc]);
    if (v > vmax) { vmax = v; p = i; }
  }
  return p;
}

// Euclidean norm with scaling (dnrm2 semantics; the summation order of the
// scaled sum is the classic LAPACK 3.x one).
template <class T>
inline T nrm2(integer n, T const* x, integer inc) {
  if (n < 1) return T(0);
  if (n == 1) return std::abs(x[0]);
  T scale = 0, ssq = 1;
  for (integer i = 0; i < n; ++i) {
    T v = x[i * inc];
    if (v != T(0)) {
      T a = std::abs(v);
      if (scale < a) { ssq = T(1) + ssq * sq(scale / a); scale = a; }
      else           { ssq += sq(a / scale); }
    }
  }
  return scale * std::sqrt(ssq);
}

// LU with partial pivoting, m x n, ipiv 1-based. Returns LAPACK info.
template <class T>
integer getrf(integer m, integer n, T* A, integer lda, integer* ipiv) {
  integer info = 0;
  integer const k = std::min(m, n);
  T const sfmin = std::numeric_limits<T>::min();
  for (integer j = 0; j < k; ++j) {
    integer const p = j + iamax(m - j, A + j + j * lda, 1);
    ipiv[j] = p + 1;
    if (A[p + j * lda] != T(0)) {
      if (p != j)
        for (integer c = 0; c < n; ++c) std::swap(A[j + c * lda], A[p + c * lda]);
      T const piv = A[j + j * lda];
      if (std::abs(piv) >= sfmin) {
        T const r = T(1) / piv;
        for (integer i = j + 1; i < m; ++i) A[i + j * lda] *= r;
      } else {
        for (integer i = j + 1; i < m; ++i) A[i + j * lda] /= piv;
      }
    } else if (info == 0) {
      info = j + 1;
    }
    for (integer c = j + 1; c < n; ++c) {
      T const t = A[j + c * lda];
      if (t != T(0))
        for (integer i = j + 1; i < m; ++i) A[i + c * lda] -= A[i + j * lda] * t;
    }
  }
  return info;
}

// forward row interchanges k1..k2 (0-based, inclusive) of an m x ncol block
template <class T>
void laswp(integer ncol, T* A, integer lda, integer k1, integer k2, integer const* ipiv) {
  for (integer i = k1; i <= k2; ++i) {
    integer const p = ipiv[i] - 1;
    if (p != i)
      for (integer c = 0; c < ncol; ++c) std::swap(A[i + c * lda], A[p + c * lda]);
  }
}

// dlarfg: generates H with H*[alpha;x] = [beta;0]
template <class T>
void larfg(integer n, T& alpha, T* x, integer incx, T& tau) {
  if (n <= 1) { tau = 0; return; }
  T const xnorm = nrm2(n - 1, x, incx);
  if (xnorm == T(0)) { tau = 0; return; }
  T beta = -std::copysign(std::hypot(alpha, xnorm), alpha);
  tau = (beta - alpha) / beta;
  T const s = T(1) / (alpha - beta);
  for (integer i = 0; i < n - 1; ++i) x[i * incx] *= s;
  alpha = beta;
}

// dlarf, SIDE='L': C(m x ncol) <- (I - tau v v^T) C, v(0) == 1 implied
template <class T>
void larf_left(integer m, integer ncol, T const* v, T tau, T* C, integer ldc) {
  if (tau == T(0)) return;
  for (integer c = 0; c < ncol; ++c) {
    T* Cc = C + c * ldc;
    T w = Cc[0];
    for (integer i = 1; i < m; ++i) w += v[i] * Cc[i];
    w *= tau;
    Cc[0] -= w;
    for (integer i = 1; i < m; ++i) Cc[i] -= v[i] * w;
  }
}

// Householder QR of m x n (dgeqr2). tau has min(m,n) entries.
template <class T>
void geqrf(integer m, integer n, T* A, integer lda, T* tau) {
  integer const k = std::min(m, n);
  std::vector<T> v(m);
  for (integer j = 0; j < k; ++j) {
    larfg(m - j, A[j + j * lda], A + std::min(j + 1, m - 1) + j * lda, 1, tau[j]);
    if (j + 1 < n) {
      v[0] = 1;
      for (integer i = 1; i < m - j; ++i) v[i] = A[j + i + j * lda];
      larf_left(m - j, n - j - 1, v.data(), tau[j], A + j + (j + 1) * lda, lda);
    }
  }
}

// C(m x ncol) <- Q^T C with Q = H_0 ... H_{k-1} stored in A (dorm2r 'L','T')
template <class T>
void ormqr_lt(integer m, integer ncol, integer k, T const* A, integer lda, T const* tau,
              T* C, integer ldc) {
  std::vector<T> v(m);
  for (integer j = 0; j < k; ++j) {
    v[0] = 1;
    for (integer i = 1; i < m - j; ++i) v[i] = A[j + i + j * lda];
    larf_left(m - j, ncol, v.data(), tau[j], C + j, ldc);
  }
}

// C(m x ncol) <- Q C (dorm2r 'L','N'): H_{k-1} applied first
template <class T>
void ormqr_ln(integer m, integer ncol, integer k, T const* A, integer lda, T const* tau,
              T* C, integer ldc) {
  std::vector<T> v(m);
  for (integer j = k - 1; j >= 0; --j) {
    v[0] = 1;
    for (integer i = 1; i < m - j; ++i) v[i] = A[j + i + j * lda];
    larf_left(m - j, ncol, v.data(), tau[j], C + j, ldc);
  }
}

// QR with column pivoting, all columns free (dgeqp3 -> dlaqp2).
// jpvt is 1-based on output.
template <class T>
void geqp3(integer m, integer n, T* A, integer lda, integer* jpvt, T* tau) {
  integer const k = std::min(m, n);
  T const tol3z = std::sqrt(std::numeric_limits<T>::epsilon());
  std::vector<T> vn1(n), vn2(n), v(m);
  for (integer j = 0; j < n; ++j) {
    jpvt[j] = j + 1;
    vn1[j] = nrm2(m, A + j * lda, 1);
    vn2[j] = vn1[j];
  }
  for (integer i = 0; i < k; ++i) {
    integer const pvt = i + iamax(n - i, vn1.data() + i, 1);
    if (pvt != i) {
      for (integer r = 0; r < m; ++r) std::swap(A[r + pvt * lda], A[r + i * lda]);
      std::swap(jpvt[pvt], jpvt[i]);
      vn1[pvt] = vn1[i];
      vn2[pvt] = vn2[i];
    }
    if (i < m - 1) larfg(m - i, A[i + i * lda], A + i + 1 + i * lda, 1, tau[i]);
    else           tau[i] = 0;  // dlarfg(1,...) gives tau = 0
    if (i + 1 < n) {
      v[0] = 1;
      for (integer r = 1; r < m - i; ++r) v[r] = A[i + r + i * lda];
      larf_left(m - i, n - i - 1, v.data(), tau[i], A + i + (i + 1) * lda, lda);
    }
    for (integer j = i + 1; j < n; ++j) {
      if (vn1[j] != T(0)) {
        T temp = T(1) - sq(std::abs(A[i + j * lda]) / vn1[j]);
        temp = std::max(temp, T(0));
        T const temp2 = temp * sq(vn1[j] / vn2[j]);
        if (temp2 <= tol3z) {
          if (i < m - 1) {
            vn1[j] = nrm2(m - i - 1, A + i + 1 + j * lda, 1);
            vn2[j] = vn1[j];
          } else {
            vn1[j] = 0;
            vn2[j] = 0;
          }
        } else {
          vn1[j] *= std::sqrt(temp);
        }
      }
    }
  }
}

// ---- BLAS-3/2 loops (reference loop order) --------------------------------

// C(m x n) = alpha*A(m x k)*B(k x n) + beta*C   (no transposes)
template <class T>
void gemm_nn(integer m, integer n, integer k, T alpha, T const* A, integer lda,
             T const* B, integer ldb, T beta, T* C, integer ldc) {
  for (integer j = 0; j < n; ++j) {
    T* Cj = C + j * ldc;
    if (beta == T(0))      for (integer i = 0; i < m; ++i) Cj[i] = 0;
    else if (beta != T(1)) for (integer i = 0; i < m; ++i) Cj[i] *= beta;
    for (integer l = 0; l < k; ++l) {
      T const t = alpha * B[l + j * ldb];
      if (t != T(0)) {
        T const* Al = A + l * lda;
        for (integer i = 0; i < m; ++i) Cj[i] += t * Al[i];
      }
    }
  }
}

// y = alpha*A*x + beta*y
template <class T>
void gemv_n(integer m, integer n, T alpha, T const* A, integer lda, T const* x, integer incx,
            T beta, T* y, integer incy) {
  if (beta != T(1))
    for (integer i = 0; i < m; ++i) y[i * incy] = (beta == T(0)) ? T(0) : beta * y[i * incy];
  for (integer j = 0; j < n; ++j) {
    T const t = alpha * x[j * incx];
    if (t != T(0)) {
      T const* Aj = A + j * lda;
      for (integer i = 0; i < m; ++i) y[i * incy] += t * Aj[i];
    }
  }
}

// B(m x ncol) <- L^{-1} B, L unit lower m x m   (trsm L,L,N,U)
template <class T>
void trsm_llnu(integer m, integer ncol, T const* L, integer ldl, T* B, integer ldb) {
  for (integer j = 0; j < ncol; ++j) {
    T* Bj = B + j * ldb;
    for (integer k = 0; k < m; ++k) {
      T const t = Bj[k];
      if (t != T(0))
        for (integer i = k + 1; i < m; ++i) Bj[i] -= t * L[i + k * ldl];
    }
  }
}

// B(m x ncol) <- U^{-1} B, U upper non-unit m x m   (trsm L,U,N,N)
template <class T>
void trsm_lunn(integer m, integer ncol, T const* U, integer ldu, T* B, integer ldb) {
  for (integer j = 0; j < ncol; ++j) {
    T* Bj = B + j * ldb;
    for (integer k = m - 1; k >= 0; --k) {
      if (Bj[k] != T(0)) {
        Bj[k] /= U[k + k * ldu];
        T const t = Bj[k];
        for (integer i = 0; i < k; ++i) Bj[i] -= t * U[i + k * ldu];
      }
    }
  }
}

// B(m x n) <- B U^{-1}, U upper non-unit n x n   (trsm R,U,N,N)
template <class T>
void trsm_runn(integer m, integer n, T const* U, integer ldu, T* B, integer ldb) {
  for (integer j = 0; j < n; ++j) {
    T* Bj = B + j * ldb;
    for (integer k = 0; k < j; ++k) {
      T const t = U[k + j * ldu];
      if (t != T(0)) {
        T const* Bk = B + k * ldb;
        for (integer i = 0; i < m; ++i) Bj[i] -= t * Bk[i];
      }
    }
    T const r = T(1) / U[j + j * ldu];
    for (integer i = 0; i < m; ++i) Bj[i] *= r;
  }
}

// x <- L^{-1} x  (trsv L,N,U, inc 1)
template <class T>
void trsv_lnu(integer n, T const* L, integer ldl, T* x) { trsm_llnu(n, 1, L, ldl, x, n); }

// x <- U^{-1} x  (trsv U,N,N, inc 1)
template <class T>
void trsv_unn(integer n, T const* U, integer ldu, T* x) { trsm_lunn(n, 1, U, ldu, x, n); }

template <class T>
void gecopy(integer m, integer n, T const* A, integer lda, T* B, integer ldb) {
  for (integer j = 0; j < n; ++j)
    for (integer i = 0; i < m; ++i) B[i + j * ldb] = A[i + j * lda];
}

template <class T>
void geadd(integer m, integer n, T const* A, integer lda, T* B, integer ldb) {
  for (integer j = 0; j < n; ++j)
    for (integer i = 0; i < m; ++i) B[i + j * ldb] += A[i + j * lda];
}

template <class T>
void gezero(integer m, integer n, T* A, integer lda) {
  for (integer j = 0; j < n; ++j)
    for (integer i = 0; i < m; ++i) A[i + j * lda] = 0;
}

}  // namespace oracle
