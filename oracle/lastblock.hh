// ORACLE — test infrastructure only. Nothing in the product links this file.
//
// Dense solvers for the bordered last block (the Nr x Nc matrix assembled in
// BABD_BorderedCR.cc:1281-1374).  The reference uses the lapack_wrapper
// classes LU / LUPQ / QR / QRP / SVD / LSS / LSY / PINV
// (BABD_BorderedCR.hxx:206-213); lapack_wrapper is an un-vendored submodule,
// so these restate the published algorithms the class names stand for:
//
//   LU   : getrf/getrs, fails on an exactly zero pivot (LAPACK info > 0)
//   LUPQ : LU with complete pivoting (max |a_ij| of the trailing block,
//          first in column-major order), fails on an exactly zero pivot
//   QR   : Householder QR (geqrf), x = R^{-1} (Q^T b)(0:Nc), needs Nr >= Nc
//   QRP  : geqp3, x = P R^{-1} (Q^T b)(0:Nc), needs Nr >= Nc
//   SVD  : A = U S V^T (one-sided Jacobi), x = V S^+ U^T b
//   LSS  : minimum-norm least squares through the SVD (dgelss semantics)
//   LSY  : minimum-norm least squares (dgelsy semantics; through the SVD here)
//   PINV : x = A^+ b
//
// The rank thresholds of SVD/LSS/LSY/PINV inside lapack_wrapper are not
// recoverable from the snapshot: here singular values below
// max(Nr,Nc) * eps * s_max count as zero (parity unpinned beyond the
// reference's closed-form tests, which only use full-rank or exactly
// rank-deficient last blocks).
#pragma once

#include "dense.hh"

namespace oracle {

template <class T>
class DenseSolver {
 public:
  enum Kind { LU = 0, LUPQ = 1, QR = 2, QRP = 3, SVD = 4, LSS = 5, LSY = 6, PINV = 7 };

  // returns false when the factorization fails
  bool factorize(Kind kind, integer m, integer n, T const* A, integer lda) {
    kind_ = kind; m_ = m; n_ = n;
    F_.assign(size_t(m) * n, T(0));
    gecopy(m, n, A, lda, F_.data(), m);
    switch (kind) {
      case LU:
        if (m != n) return false;
        ipiv_.assign(n, 0);
        return getrf(n, n, F_.data(), n, ipiv_.data()) == 0;
      case LUPQ: return factorize_lupq();
      case QR:
        if (m < n) return false;
        tau_.assign(n, T(0));
        geqrf(m, n, F_.data(), m, tau_.data());
        return nonzero_diag();
      case QRP:
        if (m < n) return false;
        tau_.assign(n, T(0));
        jpvt_.assign(n, 0);
        geqp3(m, n, F_.data(), m, jpvt_.data(), tau_.data());
        return nonzero_diag();
      case SVD: case LSS: case LSY: case PINV:
        return factorize_svd();
    }
    return false;
  }

  // b has max(m,n) entries: m on input, n on output
  bool solve(T* b) const {
    switch (kind_) {
      case LU:
        laswp(1, b, n_, 0, n_ - 1, ipiv_.data());
        trsv_lnu(n_, F_.data(), n_, b);
        trsv_unn(n_, F_.data(), n_, b);
        return true;
      case LUPQ: {
        for (integer i = 0; i < n_; ++i) if (P_[i] != i) std::swap(b[i], b[P_[i]]);
        trsv_lnu(n_, F_.data(), n_, b);
        trsv_unn(n_, F_.data(), n_, b);
        for (integer i = n_ - 1; i >= 0; --i) if (Q_[i] != i) std::swap(b[i], b[Q_[i]]);
        return true;
      }
      case QR:
        ormqr_lt(m_, 1, n_, F_.data(), m_, tau_.data(), b, m_);
        trsv_unn(n_, F_.data(), m_, b);
        return true;
      case QRP: {
        ormqr_lt(m_, 1, n_, F_.data(), m_, tau_.data(), b, m_);
        trsv_unn(n_, F_.data(), m_, b);
        std::vector<T> z(b, b + n_);
        for (integer i = 0; i < n_; ++i) b[jpvt_[i] - 1] = z[i];
        return true;
      }
      default: {
        // x = V diag(1/s) U^T b over the retained singular triplets
        integer const k = std::min(m_, n_);
        std::vector<T> c(k, T(0)), x(n_, T(0));
        for (integer l = 0; l < rank_; ++l) {
          T acc = 0;
          for (integer i = 0; i < m_; ++i) acc += U_[i + l * m_] * b[i];
          c[l] = acc / S_[l];
        }
        for (integer l = 0; l < rank_; ++l)
          for (integer i = 0; i < n_; ++i) x[i] += V_[i + l * n_] * c[l];
        for (integer i = 0; i < n_; ++i) b[i] = x[i];
        return true;
      }
    }
  }

  integer rank() const { return rank_; }

 private:
  bool nonzero_diag() const {
    for (integer i = 0; i < n_; ++i) if (F_[i + i * m_] == T(0)) return false;
    return true;
  }

  bool factorize_lupq() {
    if (m_ != n_) return false;
    integer const n = n_;
    P_.assign(n, 0); Q_.assign(n, 0);
    T* A = F_.data();
    for (integer k = 0; k < n; ++k) {
      integer pi = k, pj = k;
      T amax = -1;
      for (integer j = k; j < n; ++j)
        for (integer i = k; i < n; ++i) {
          T v = std::abs(A[i + j * n]);
          if (v > amax) { amax = v; pi = i; pj = j; }
        }
      P_[k] = pi; Q_[k] = pj;
      if (pi != k) for (integer c = 0; c < n; ++c) std::swap(A[k + c * n], A[pi + c * n]);
      if (pj != k) for (integer r = 0; r < n; ++r) std::swap(A[r + k * n], A[r + pj * n]);
      T const piv = A[k + k * n];
      if (piv == T(0)) return false;
      T const r = T(1) / piv;
      for (integer i = k + 1; i < n; ++i) A[i + k * n] *= r;
      for (integer c = k + 1; c < n; ++c) {
        T const t = A[k + c * n];
        if (t != T(0))
          for (integer i = k + 1; i < n; ++i) A[i + c * n] -= A[i + k * n] * t;
      }
    }
    return true;
  }

  // one-sided Jacobi SVD of the m x n matrix in F_
  bool factorize_svd() {
    bool const tr = m_ < n_;
    integer const rows = tr ? n_ : m_;
    integer const cols = tr ? m_ : n_;
    std::vector<T> W(size_t(rows) * cols), Vj(size_t(cols) * cols, T(0));
    for (integer j = 0; j < cols; ++j)
      for (integer i = 0; i < rows; ++i)
        W[i + j * rows] = tr ? F_[j + i * m_] : F_[i + j * m_];
    for (integer j = 0; j < cols; ++j) Vj[j + j * cols] = 1;
    T const eps = std::numeric_limits<T>::epsilon();
    for (int sweep = 0; sweep < 80; ++sweep) {
      T off = 0;
      for (integer p = 0; p < cols - 1; ++p)
        for (integer q = p + 1; q < cols; ++q) {
          T a = 0, b = 0, g = 0;
          for (integer i = 0; i < rows; ++i) {
            T const wp = W[i + p * rows], wq = W[i + q * rows];
            a += wp * wp; b += wq * wq; g += wp * wq;
          }
          if (g == T(0)) continue;
          T const rel = std::abs(g) / std::sqrt(a * b);
          off = std::max(off, rel);
          if (rel <= eps) continue;
          T const zeta = (b - a) / (2 * g);
          T const t = std::copysign(T(1), zeta) / (std::abs(zeta) + std::sqrt(T(1) + zeta * zeta));
          T const c = T(1) / std::sqrt(T(1) + t * t), s = c * t;
          for (integer i = 0; i < rows; ++i) {
            T const wp = W[i + p * rows], wq = W[i + q * rows];
            W[i + p * rows] = c * wp - s * wq;
            W[i + q * rows] = s * wp + c * wq;
          }
          for (integer i = 0; i < cols; ++i) {
            T const vp = Vj[i + p * cols], vq = Vj[i + q * cols];
            Vj[i + p * cols] = c * vp - s * vq;
            Vj[i + q * cols] = s * vp + c * vq;
          }
        }
      if (off <= eps) break;
    }
    // singular values = column norms, sorted descending
    std::vector<T> s(cols);
    std::vector<integer> ord(cols);
    for (integer j = 0; j < cols; ++j) { s[j] = nrm2(rows, W.data() + j * rows, 1); ord[j] = j; }
    std::stable_sort(ord.begin(), ord.end(), [&](integer a, integer b) { return s[a] > s[b]; });
    T const smax = cols > 0 ? s[ord[0]] : T(0);
    T const tol = T(std::max(m_, n_)) * eps * smax;
    rank_ = 0;
    S_.assign(cols, T(0));
    // left vectors of W are W(:,j)/s_j, right vectors are Vj(:,j);
    // for the transposed problem the roles swap.
    U_.assign(size_t(m_) * cols, T(0));
    V_.assign(size_t(n_) * cols, T(0));
    for (integer l = 0; l < cols; ++l) {
      integer const j = ord[l];
      S_[l] = s[j];
      if (s[j] > tol && s[j] > T(0)) ++rank_;
      for (integer i = 0; i < rows; ++i) {
        T const u = s[j] > T(0) ? W[i + j * rows] / s[j] : T(0);
        if (tr) V_[i + l * n_] = u; else U_[i + l * m_] = u;
      }
      for (integer i = 0; i < cols; ++i) {
        T const v = Vj[i + j * cols];
        if (tr) U_[i + l * m_] = v; else V_[i + l * n_] = v;
      }
    }
    return rank_ > 0;
  }

  Kind kind_{LU};
  integer m_{0}, n_{0}, rank_{0};
  std::vector<T> F_, tau_, S_, U_, V_;
  std::vector<integer> ipiv_, jpvt_, P_, Q_;
};

}  // namespace oracle
