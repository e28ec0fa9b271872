// Host side of alglin_amd::BorderedCR<T> (see BorderedCR.hh).
#include "BorderedCR.hh"

#include "bcr_launch.hh"

#include <cmath>
#include <sstream>
#include <stdexcept>
#include <type_traits>

namespace alglin_amd {

namespace {

template <typename T>
void gecopy(integer m, integer n, T const* A, integer lda, T* B, integer ldb) {
  for (integer j = 0; j < n; ++j)
    for (integer i = 0; i < m; ++i) B[i + size_t(j) * ldb] = A[i + size_t(j) * lda];
}

template <typename T>
void geadd(integer m, integer n, T const* A, integer lda, T* B, integer ldb) {
  for (integer j = 0; j < n; ++j)
    for (integer i = 0; i < m; ++i) B[i + size_t(j) * ldb] += A[i + size_t(j) * lda];
}

template <typename T>
void gemv(integer m, integer n, T alpha, T const* A, integer lda, T const* x, T* y) {
  for (integer j = 0; j < n; ++j) {
    T const t = alpha * x[j];
    for (integer i = 0; i < m; ++i) y[i] += t * A[i + size_t(j) * lda];
  }
}

}  // namespace

template <typename T>
BorderedCR<T>::BorderedCR(void*, integer) : eng_(new bcr::Engine<T, T>()) {}

template <typename T>
BorderedCR<T>::~BorderedCR() = default;

template <typename T>
void BorderedCR<T>::fail(std::string const& msg) const {
  last_error_ = msg;
  throw std::runtime_error(msg);
}

// cc:42-170 (the range partition of the host thread pool has no GPU meaning)
template <typename T>
void BorderedCR<T>::allocate(integer N, integer n, integer qr, integer qx, integer nr, integer nx) {
  if (N < 1 || n < 1 || qr < 0 || qx < 0 || nr < 0 || nx < 0)
    fail("BorderedCR::allocate, bad dimensions");
  bcr::Dims d;
  d.N = N; d.n = n; d.qr = qr; d.qx = qx; d.nr = nr; d.nx = nx;
  if (d == d_ && !D_.empty()) return;  // cc:55-61
  d_ = d;
  D_.assign(size_t(N) * n * n, T(0));
  E_.assign(size_t(N) * n * n, T(0));
  B_.assign(size_t(N) * n * nx, T(0));
  C_.assign(size_t(N + 1) * nr * n, T(0));
  Cq_.assign(size_t(nr) * qx, T(0));
  F_.assign(size_t(nr) * nx, T(0));
  H_.assign(size_t(n + qr) * d.Nc(), T(0));
  try {
    eng_->allocate(d_);
    if (mix_) mix_->allocate(d_);
  } catch (std::exception const& e) {
    last_error_ = e.what();
    throw;
  }
  factorized_ = false;
}

template <typename T>
void BorderedCR<T>::dup(BorderedCR const& M) {
  allocate(M.d_.N, M.d_.n, M.d_.qr, M.d_.qx, M.d_.nr, M.d_.nx);
  D_ = M.D_; E_ = M.E_; B_ = M.B_; C_ = M.C_; Cq_ = M.Cq_; F_ = M.F_; H_ = M.H_;
}

template <typename T>
void BorderedCR<T>::select(std::string_view s) {
  if (s == "LU") sel_ = Choice::LU;
  else if (s == "QR") sel_ = Choice::QR;
  else if (s == "QRP") sel_ = Choice::QRP;
  else throw std::invalid_argument("BorderedCR::select( choice='" + std::string(s) +
                                   "' )\nunknown/unsupported factorization type ('LU','QR','QRP')");
}

template <typename T>
void BorderedCR<T>::select_last(std::string_view s) {
  static char const* names[] = {"LU", "LUPQ", "QR", "QRP", "SVD", "LSS", "LSY", "PINV"};
  for (int i = 0; i < 8; ++i)
    if (s == names[i]) { last_ = LastChoice(i); return; }
  throw std::invalid_argument("BorderedCR::select_last( choice='" + std::string(s) +
                              "' )\nunknown/unsupported factorization type ('LU','LUPQ','QR','QRP','SVD','LSS','LSY','PINV')");
}

template <typename T>
void BorderedCR<T>::select_last2(std::string_view s) {
  static char const* names[] = {"NONE", "SVD", "LSS", "LSY", "PINV"};
  for (int i = 0; i < 5; ++i)
    if (s == names[i]) { last2_ = LastChoice2(i); return; }
  throw std::invalid_argument("BorderedCR::select_last2( choice='" + std::string(s) +
                              "' )\nunknown/unsupported factorization type ('NONE','SVD','LSS','LSY','PINV')");
}

static char const* kChoice[] = {"CyclicReduction+LU", "CyclicReduction+QR", "CyclicReduction+QRP"};
static char const* kLast[] = {"LastBlock LU", "LastBlock LUPQ", "LastBlock QR", "LastBlock QRP",
                              "LastBlock SVD", "LastBlock LSS", "LastBlock LSY", "LastBlock PINV"};
static char const* kLast2[] = {"LastBlock NONE", "LastBlock SVD", "LastBlock LSS", "LastBlock LSY",
                               "LastBlock PINV"};

template <typename T>
std::string BorderedCR<T>::info_algo() const {
  return std::string(kChoice[int(sel_)]) + " and " + kLast[int(last_)] + " and " + kLast2[int(last2_)];
}

// cc:208-235; the reference prints thread_count()/nblk (host pool); the GPU
// solver runs one cyclic reduction over all blocks, i.e. nblk:1.
template <typename T>
std::string BorderedCR<T>::info(std::string_view ind) const {
  std::ostringstream s;
  std::string const I(ind);
  s << std::boolalpha << I << "rows   = " << nrows() << "\n" << I << "cols   = " << ncols() << "\n"
    << I << "nblock = " << d_.N << "\n" << I << "n      = " << d_.n << "\n" << I << "qr     = " << d_.qr << "\n"
    << I << "qx     = " << d_.qx << "\n" << I << "nr     = " << d_.nr << "\n" << I << "nx     = " << d_.nx << "\n"
    << I << "treads = 1/nblk:1/factorize:" << fac_thread_ << "/solve:" << sol_thread_ << "\n"
    << I << "intern = " << kChoice[int(sel_)] << "\n" << I << "last   = " << kLast[int(last_)] << "\n"
    << I << "last2  = " << kLast2[int(last2_)] << "\n" << I << "device = MI355X (gfx950)\n";
  return s.str();
}

template <typename T> void BorderedCR<T>::zero_D() { std::fill(D_.begin(), D_.end(), T(0)); }
template <typename T> void BorderedCR<T>::zero_E() { std::fill(E_.begin(), E_.end(), T(0)); }
template <typename T> void BorderedCR<T>::zero_B() { std::fill(B_.begin(), B_.end(), T(0)); }
template <typename T> void BorderedCR<T>::zero_F() { std::fill(F_.begin(), F_.end(), T(0)); }
template <typename T> void BorderedCR<T>::zero_H() { std::fill(H_.begin(), H_.end(), T(0)); }
template <typename T> void BorderedCR<T>::zero_C() { std::fill(C_.begin(), C_.end(), T(0)); }
template <typename T> void BorderedCR<T>::zero_Cq() { std::fill(Cq_.begin(), Cq_.end(), T(0)); }

template <typename T>
void BorderedCR<T>::load_B(integer k, T const B[], integer ld) {
  gecopy(d_.n, d_.nx, B, ld, &B_[size_t(k) * d_.n * d_.nx], d_.n);
}
template <typename T>
void BorderedCR<T>::add_to_B(integer k, T const B[], integer ld) {
  geadd(d_.n, d_.nx, B, ld, &B_[size_t(k) * d_.n * d_.nx], d_.n);
}
template <typename T>
void BorderedCR<T>::load_C(integer k, T const C[], integer ld) {
  gecopy(d_.nr, d_.n, C, ld, &C_[size_t(k) * d_.nr * d_.n], d_.nr);
}
template <typename T>
void BorderedCR<T>::add_to_C(integer k, T const C[], integer ld) {
  if (ld < d_.nr) fail("BorderedCR::add_to_C( " + std::to_string(k) + ", C, ldC = " + std::to_string(ld) + " ) bad ldC\n");
  geadd(d_.nr, d_.n, C, ld, &C_[size_t(k) * d_.nr * d_.n], d_.nr);
}
template <typename T>
void BorderedCR<T>::add_to_C2(integer k, T const C[], integer ld) {
  if (ld < d_.nr) fail("BorderedCR::add_to_C2( " + std::to_string(k) + ", C, ldC = " + std::to_string(ld) + " ) bad ldC\n");
  geadd(d_.nr, 2 * d_.n, C, ld, &C_[size_t(k) * d_.nr * d_.n], d_.nr);
}
template <typename T>
void BorderedCR<T>::add_to_C2F(integer k, T const C2F[], integer ld) {
  if (ld < d_.nr) fail("BorderedCR::add_to_C2F( " + std::to_string(k) + ", C2F, ldC = " + std::to_string(ld) + " ) bad ldC\n");
  geadd(d_.nr, 2 * d_.n, C2F, ld, &C_[size_t(k) * d_.nr * d_.n], d_.nr);
  geadd(d_.nr, d_.nx, C2F + size_t(ld) * 2 * d_.n, ld, F_.data(), d_.nr);
}
template <typename T>
void BorderedCR<T>::load_D(integer k, T const D[], integer ld) {
  gecopy(d_.n, d_.n, D, ld, &D_[size_t(k) * d_.n * d_.n], d_.n);
}
template <typename T>
void BorderedCR<T>::load_E(integer k, T const E[], integer ld) {
  gecopy(d_.n, d_.n, E, ld, &E_[size_t(k) * d_.n * d_.n], d_.n);
}
template <typename T>
void BorderedCR<T>::load_DE(integer k, T const DE[], integer ld) {
  load_D(k, DE, ld);
  load_E(k, DE + size_t(d_.n) * ld, ld);
}
template <typename T>
void BorderedCR<T>::load_DEB(integer k, T const DEB[], integer ld) {
  load_D(k, DEB, ld);
  load_E(k, DEB + size_t(d_.n) * ld, ld);
  load_B(k, DEB + size_t(2 * d_.n) * ld, ld);
}
template <typename T>
void BorderedCR<T>::load_F(T const F[], integer ld) { gecopy(d_.nr, d_.nx, F, ld, F_.data(), d_.nr); }
template <typename T>
void BorderedCR<T>::add_to_F(T const F[], integer ld) { geadd(d_.nr, d_.nx, F, ld, F_.data(), d_.nr); }
template <typename T>
void BorderedCR<T>::load_Cq(T const Cq[], integer ld) { gecopy(d_.nr, d_.qx, Cq, ld, Cq_.data(), d_.nr); }
template <typename T>
void BorderedCR<T>::load_CqF(T const CqF[], integer ld) {
  gecopy(d_.nr, d_.qx, CqF, ld, Cq_.data(), d_.nr);
  gecopy(d_.nr, d_.nx, CqF + size_t(d_.qx) * ld, ld, F_.data(), d_.nr);
}
template <typename T>
void BorderedCR<T>::load_bottom(T const H0[], integer ld0, T const HN[], integer ldN, T const Hq[], integer ldQ,
                                T const Hp[], integer ldP) {
  integer const m = d_.n + d_.qr;
  T* H = H_.data();
  gecopy(m, d_.n, H0, ld0, H, m); H += size_t(m) * d_.n;
  gecopy(m, d_.n, HN, ldN, H, m); H += size_t(m) * d_.n;
  gecopy(m, d_.qx, Hq, ldQ, H, m); H += size_t(m) * d_.qx;
  gecopy(m, d_.nx, Hp, ldP, H, m);
}
template <typename T>
void BorderedCR<T>::load_bottom(T const H0Nqp[], integer ldH) {
  gecopy(d_.n + d_.qr, d_.Nc(), H0Nqp, ldH, H_.data(), d_.n + d_.qr);
}
template <typename T>
void BorderedCR<T>::load_bottom2(T const C0[], integer ld0, T const CN[], integer ldN, T const Cq[], integer ldCq,
                                 T const F[], integer ldF) {
  gecopy(d_.nr, d_.n, C0, ld0, C_.data(), d_.nr);
  gecopy(d_.nr, d_.n, CN, ldN, &C_[size_t(d_.N) * d_.nr * d_.n], d_.nr);
  gecopy(d_.nr, d_.qx, Cq, ldCq, Cq_.data(), d_.nr);
  gecopy(d_.nr, d_.nx, F, ldF, F_.data(), d_.nr);
}

template <typename T>
bool BorderedCR<T>::factorize() {
  factorized_ = false;
  try {
    if (debug_) check_matrix();
    int32_t st = 0;
    device_bound_ = false;   // the host blocks become the operator
    if (mixed_) {
      if constexpr (std::is_same_v<T, double>) {
        auto& e = mix_->engine();
        e.set_method(int(sel_));
        e.set_last_method(int(last_));
        e.set_last2(last2_ != LastChoice2::NONE);
        e.upload(D_.data(), E_.data(), B_.data(), C_.data(), H_.data(), Cq_.data(), F_.data());
        e.factorize_async();
        st = e.status_sync();
      }
    } else {
      eng_->set_method(int(sel_));
      eng_->set_last_method(int(last_));
      eng_->set_last2(last2_ != LastChoice2::NONE);
      eng_->upload(D_.data(), E_.data(), B_.data(), C_.data(), H_.data(), Cq_.data(), F_.data());
      eng_->factorize_async();
      st = eng_->status_sync();
    }
    if (st & 1) {
      last_error_ = "BorderedCR::buildT, zero pivot in a cyclic-reduction elimination";
      return false;
    }
    if (st & 2) {
      last_error_ = "BorderedCR<t_Value>::load_and_factorize_last failed\n";
      return false;
    }
  } catch (std::exception const& e) {
    last_error_ = e.what();
    return false;
  }
  factorized_ = true;
  return true;
}

template <typename T>
bool BorderedCR<T>::solve(T x[]) {
  return solve(1, x, std::max(nrows(), ncols()));
}

template <typename T>
bool BorderedCR<T>::solve(integer nrhs, T X[], integer ld) {
  if (!factorized_) {
    last_error_ = "BorderedCR::solve, matrix not factorized";
    return false;
  }
  try {
    const integer L = std::max(nrows(), ncols());
    if (ld < L) throw std::invalid_argument("BorderedCR::solve, ldRhs < max(nrows,ncols)");
    xbuf_.resize(size_t(L) * nrhs);
    hipStream_t st = eng_->stream();
    if constexpr (std::is_same_v<T, double>)
      if (mixed_) st = mix_->engine().stream();
    bcr::hip_check(hipMemcpy2DAsync(xbuf_.p, size_t(L) * sizeof(T), X, size_t(ld) * sizeof(T), size_t(L) * sizeof(T),
                                    nrhs, hipMemcpyHostToDevice, st), "H2D x");
    solve_device(xbuf_.p, nrhs, L, false);
    bcr::hip_check(hipMemcpy2DAsync(X, size_t(ld) * sizeof(T), xbuf_.p, size_t(L) * sizeof(T), size_t(L) * sizeof(T),
                                    nrhs, hipMemcpyDeviceToHost, st), "D2H x");
    bcr::hip_check(hipStreamSynchronize(st), "solve sync");
    if (mixed_ && report_.status != 0 && report_.status != 4) {
      last_error_ = "BorderedCR::solve, iterative refinement did not reach the requested residual";
      return false;
    }
  } catch (std::exception const& e) {
    last_error_ = e.what();
    return false;
  }
  return true;
}

template <typename T>
void BorderedCR<T>::select_precision(std::string_view mode, double tol, int max_iter) {
  if (mode == "FP64" || mode == "NATIVE") {
    mixed_ = false;
  } else if (mode == "MIXED") {
    if constexpr (!std::is_same_v<T, double>)
      throw std::invalid_argument("BorderedCR::select_precision: MIXED needs the double solver");
    if (!mix_) {
      mix_.reset(new bcr::MixedEngine());
      mix_->engine().share_stream(eng_->stream());   // one stream per object
      if (d_.N > 0) mix_->allocate(d_);
    }
    mixed_ = true;
  } else {
    throw std::invalid_argument("BorderedCR::select_precision( mode='" + std::string(mode) +
                                "' ) unknown mode ('FP64','MIXED')");
  }
  ir_tol_ = tol;
  ir_max_ = max_iter;
  factorized_ = false;
}

template <typename T>
void BorderedCR<T>::bind_device(T const* D, T const* E, T const* B, T const* C, T const* H, T const* Cq, T const* F) {
  device_bound_ = true;
  eng_->bind_device(D, E, B, C, H, Cq, F);
  if constexpr (std::is_same_v<T, double>)
    if (mix_) mix_->engine().bind_device(D, E, B, C, H, Cq, F);
}

template <typename T>
int32_t BorderedCR<T>::factorize_device(bool sync) {
  int32_t st = 0;
  if (mixed_) {
    if constexpr (std::is_same_v<T, double>) {
      auto& e = mix_->engine();
      e.set_method(int(sel_));
      e.set_last_method(int(last_));
      e.set_last2(last2_ != LastChoice2::NONE);
      e.factorize_async();
      if (sync) st = e.status_sync();
    }
  } else {
    eng_->set_method(int(sel_));
    eng_->set_last_method(int(last_));
    eng_->set_last2(last2_ != LastChoice2::NONE);
    eng_->factorize_async();
    if (sync) st = eng_->status_sync();
  }
  st &= 3;   // bit 3 only records that the second last-block solver was used
  factorized_ = st == 0;
  return st;
}

// in-place solve of device columns; in MIXED mode the columns are refined
// together, and when refinement stalls the FP64 QRP path solves them again
template <typename T>
void BorderedCR<T>::solve_device(T* X, integer nrhs, int64_t ld, bool sync) {
  if (mixed_) {
    if constexpr (std::is_same_v<T, double>) {
      report_ = mix_->solve(X, nrhs, ld, ir_tol_, ir_max_);
      if (report_.status != 0 && fallback_) fallback_qrp(X, nrhs, ld);
    }
    return;
  }
  eng_->template solve_async<T>(X, nrhs, ld);
  if (sync) bcr::hip_check(hipStreamSynchronize(eng_->stream()), "solve sync");
}

// FP64 cyclic reduction with column-pivoted QR on the pristine operator,
// applied to the right-hand sides the refinement kept
template <typename T>
void BorderedCR<T>::fallback_qrp(T* X, integer nrhs, int64_t ld) {
  if constexpr (std::is_same_v<T, double>) {
    if (!device_bound_) eng_->upload(D_.data(), E_.data(), B_.data(), C_.data(), H_.data(), Cq_.data(), F_.data());
    eng_->set_method(int(Choice::QRP));
    eng_->set_last_method(int(last_));
    eng_->set_last2(last2_ != LastChoice2::NONE);
    eng_->factorize_async();
    const int32_t st = eng_->status_sync();
    eng_->set_method(int(sel_));
    if (st != 0) {
      report_.status = 3;
      return;
    }
    const int64_t L = mix_->rhs_ld();
    bcr::hip_check(hipMemcpy2DAsync(X, size_t(ld) * sizeof(double), mix_->rhs(), size_t(L) * sizeof(double),
                                    size_t(L) * sizeof(double), nrhs, hipMemcpyDeviceToDevice, eng_->stream()),
                   "fallback rhs");
    eng_->template solve_async<double>(X, nrhs, ld);
    // FP64 refinement on the FP64 QRP factors (they lose digits on nearly
    // singular pair columns; the sweeps recover them)
    double rel = mix_->max_rel_residual(X, nrhs, ld, mix_->rhs(), L);
    const int64_t nrows = d_.nrows();
    for (int it = 0; it < ir_max_ && rel > ir_tol_ && rel == rel; ++it) {
      double* R = mix_->residuals();
      if (L > nrows)
        for (integer c = 0; c < nrhs; ++c)
          bcr::hip_check(bcr::launch_convert<double, double>(nullptr, R + size_t(c) * L + nrows, L - nrows,
                                                              eng_->stream()),
                         "pad");
      eng_->template solve_async<double>(R, nrhs, L);
      for (integer c = 0; c < nrhs; ++c)
        bcr::hip_check(bcr::launch_axpy(1.0, R + size_t(c) * L, X + size_t(c) * ld, d_.ncols(), eng_->stream()),
                       "x += d");
      const double prev = rel;
      rel = mix_->max_rel_residual(X, nrhs, ld, mix_->rhs(), L);
      if (!(rel < 0.9 * prev)) break;
    }
    report_.rel_residual = rel;
    report_.status = 4;
  }
}

template <typename T>
bool BorderedCR<T>::t_solve(T[]) const {
  fail("BorderedCR::t_solve() not defined\n");
  return false;
}

template <typename T>
bool BorderedCR<T>::t_solve(integer, T[], integer) const {
  fail("BorderedCR::t_solve() not defined\n");
  return false;
}

template <typename T>
void BorderedCR<T>::Mv(T const x[], T res[]) const {
  for (integer i = 0; i < nrows(); ++i) res[i] = T(0);
  add_Mv(T(1), x, res);
}

template <typename T>
void BorderedCR<T>::add_Mv(T alpha, T const x[], T res[]) const {
  const integer N = d_.N, n = d_.n, nr = d_.nr, nx = d_.nx, qx = d_.qx;
  T const* xe = x + size_t(N) * n;
  T const* xq = xe + n;
  T const* xb = xq + qx;
  T* yy = res;
  for (integer k = 0; k < N; ++k) {
    gemv(n, n, alpha, &D_[size_t(k) * n * n], n, x + size_t(k) * n, yy);
    gemv(n, n, alpha, &E_[size_t(k) * n * n], n, x + size_t(k + 1) * n, yy);
    gemv(n, nx, alpha, &B_[size_t(k) * n * nx], n, xb, yy);
    yy += n;
  }
  integer const m = n + d_.qr;
  T const* H = H_.data();
  gemv(m, n, alpha, H, m, x, yy); H += size_t(m) * n;
  gemv(m, n, alpha, H, m, xe, yy); H += size_t(m) * n;
  gemv(m, qx, alpha, H, m, xq, yy); H += size_t(m) * qx;
  gemv(m, nx, alpha, H, m, xb, yy);
  if (nr > 0) {
    yy += m;
    gemv(nr, nx, alpha, F_.data(), nr, xb, yy);
    for (integer k = 0; k <= N; ++k) gemv(nr, n, alpha, &C_[size_t(k) * nr * n], nr, x + size_t(k) * n, yy);
    gemv(nr, qx, alpha, Cq_.data(), nr, x + size_t(N + 1) * n, yy);
  }
}

template <typename T>
integer BorderedCR<T>::sparse_nnz() const {
  const integer N = d_.N, n = d_.n, nr = d_.nr, nx = d_.nx, qr = d_.qr, qx = d_.qx;
  return n * N * (2 * n + nx + nr) + nr * (n + qx + nx) + (n + qr) * (2 * n + qx + nx);
}

// cc:2178-2378: D,E,B per block row, then H, then C_0..C_N, Cq, F
template <typename T>
void BorderedCR<T>::sparse_pattern(integer I[], integer J[], integer offs) const {
  const integer N = d_.N, n = d_.n, nr = d_.nr, nx = d_.nx, qr = d_.qr, qx = d_.qx;
  integer kk = 0;
  auto put = [&](integer i, integer j) { I[kk] = i + offs; J[kk] = j + offs; ++kk; };
  for (integer k = 0; k < N; ++k) {
    for (integer ij = 0; ij < n * n; ++ij) put(k * n + ij % n, k * n + ij / n);
    for (integer ij = 0; ij < n * n; ++ij) put(k * n + ij % n, (k + 1) * n + ij / n);
    for (integer ij = 0; ij < n * nx; ++ij) put(k * n + ij % n, (N + 1) * n + qx + ij / n);
  }
  const integer nqr = n + qr;
  for (integer ij = 0; ij < nqr * d_.Nc(); ++ij) {
    integer j = ij / nqr;
    if (j >= n) j += N * n - n;
    put(N * n + ij % nqr, j);
  }
  const integer i0 = (N + 1) * n + qr;
  for (integer k = 0; k <= N; ++k)
    for (integer ij = 0; ij < nr * n; ++ij) put(i0 + ij % nr, k * n + ij / nr);
  for (integer ij = 0; ij < nr * qx; ++ij) put(i0 + ij % nr, (N + 1) * n + ij / nr);
  for (integer ij = 0; ij < nr * nx; ++ij) put(i0 + ij % nr, (N + 1) * n + qx + ij / nr);
  if (kk != sparse_nnz()) fail("BorderedCR::sparse_pattern( I, J, offs ), inserted values count mismatch\n");
}

template <typename T>
void BorderedCR<T>::sparse_values(T V[]) const {
  const integer N = d_.N, n = d_.n, nr = d_.nr, nx = d_.nx;
  size_t kk = 0;
  for (integer k = 0; k < N; ++k) {
    for (integer ij = 0; ij < n * n; ++ij) V[kk++] = D_[size_t(k) * n * n + ij];
    for (integer ij = 0; ij < n * n; ++ij) V[kk++] = E_[size_t(k) * n * n + ij];
    for (integer ij = 0; ij < n * nx; ++ij) V[kk++] = B_[size_t(k) * n * nx + ij];
  }
  for (T v : H_) V[kk++] = v;
  for (integer k = 0; k <= N; ++k)
    for (integer ij = 0; ij < nr * n; ++ij) V[kk++] = C_[size_t(k) * nr * n + ij];
  for (T v : Cq_) V[kk++] = v;
  for (T v : F_) V[kk++] = v;
}

// cc:2417-2482
template <typename T>
void BorderedCR<T>::sparse_load(T const V[], integer const Ir[], integer r_offs, integer const Jc[], integer c_offs,
                                integer nnz, bool do_error) {
  const integer N = d_.N, n = d_.n, nr = d_.nr, qr = d_.qr, qx = d_.qx, nx = d_.nx;
  const integer rH = n * N, rC = rH + n + qr, nrow = rC + nr;
  const integer cCq = rH + n, cF = cCq + qx, ncol = cF + nx;
  fill_zero();
  for (integer k = 0; k < nnz; ++k) {
    const integer i = Ir[k] - r_offs, j = Jc[k] - c_offs;
    const T v = V[k];
    const integer ib = i / n, jb = j / n;
    bool ok = true;
    if (i < rH) {
      if (j < cCq) {
        if (ib == jb) D(ib, i % n, j % n) = v;
        else if (ib + 1 == jb) E(ib, i % n, j % n) = v;
        else ok = false;
      } else if (j < cF) ok = false;
      else if (j < ncol) B(ib, i % n, j - cF) = v;
      else ok = false;
    } else if (i < rC) {
      if (j < n) H(i - rH, j) = v;
      else if (j < rH) ok = false;
      else if (j < ncol) H(i - rH, j - rH + n) = v;
      else ok = false;
    } else if (i < nrow) {
      if (j < cCq) C(jb, i - rC, j % n) = v;
      else if (j < cF) Cq(i - rC, j - cCq) = v;
      else if (j < ncol) F(i - rC, j - cF) = v;
      else ok = false;
    } else {
      ok = false;
    }
    if (!ok) {
      last_error_ = "in BorderedCR<t_Value>::sparse_load, indices (i,j) = ( " + std::to_string(Ir[k]) + ", " +
                    std::to_string(Jc[k]) + ") out of pattern!\n";
      if (do_error) throw std::runtime_error(last_error_);
    }
  }
}

template <typename T>
void BorderedCR<T>::check_matrix() {
  auto scan = [&](std::vector<T> const& v, char const* name) {
    for (size_t i = 0; i < v.size(); ++i)
      if (!std::isfinite(double(v[i])))
        fail(std::string("BorderedCR::check_matrix, found ") + std::to_string(double(v[i])) + " in matrix " + name +
             " at linear position " + std::to_string(i) + "\n");
  };
  scan(H_, "H0Hqp"); scan(B_, "B"); scan(C_, "C"); scan(Cq_, "Cq"); scan(D_, "D"); scan(E_, "E"); scan(F_, "F");
}

template class BorderedCR<double>;
template class BorderedCR<float>;

}  // namespace alglin_amd
