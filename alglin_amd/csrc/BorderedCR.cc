// Host side of alglin_amd::BorderedCR<T> (include/alglin_amd/BorderedCR.hh).
#include "alglin_amd/BorderedCR.hh"

#include <algorithm>
#include <charconv>
#include <cmath>
#include <sstream>
#include <stdexcept>
#include <type_traits>

#include "bcr_gpu.hh"
#include "bcr_launch.hh"

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

// shortest round-trip decimal form (what fmt's "{}" prints for a float)
template <typename T>
std::string num(T v) {
  char buf[64];
  auto r = std::to_chars(buf, buf + sizeof(buf), v);
  return std::string(buf, r.ptr);
}

}  // namespace

template <typename T>
BorderedCR<T>::BorderedCR(Utils::ThreadPoolBase* TP, integer npb)
    : max_pb_(TP == nullptr ? 1 : std::max<integer>(1, npb)) {}

template <typename T>
BorderedCR<T>::~BorderedCR() = default;

template <typename T>
typename BorderedCR<T>::Gpu& BorderedCR<T>::gpu() const {
  if (!gpu_) {
    gpu_.reset(new Gpu());
    gpu_->want_P = max_pb_;
  }
  gpu_->want.N = d_N_; gpu_->want.n = d_n_; gpu_->want.qr = d_qr_;
  gpu_->want.qx = d_qx_; gpu_->want.nr = d_nr_; gpu_->want.nx = d_nx_;
  return *gpu_;
}

template <typename T>
void BorderedCR<T>::fail(std::string const& msg) const {
  last_error_ = msg;
  throw std::runtime_error(msg);
}

template <typename T>
void BorderedCR<T>::check_ld(char const* what, integer nbl, integer ld, integer rows) const {
  if (ld < std::max<integer>(1, rows))
    fail(std::string("BorderedCR::") + what + "( " + std::to_string(nbl) + ", ..., ld = " + std::to_string(ld) +
         " ) bad leading dimension, expected >= " + std::to_string(rows) + "\n");
}

template <typename T>
void BorderedCR<T>::check_dims(char const* what, MatW const& M, integer nr, integer nc) const {
  if (M.nrows() != nr || M.ncols() != nc)
    fail(std::string("BorderedCR::") + what + ", bad dimension size = " + std::to_string(M.nrows()) + " x " +
         std::to_string(M.ncols()) + " expected " + std::to_string(nr) + " x " + std::to_string(nc) + "\n");
  if (M.ldim() < std::max<integer>(1, nr))
    fail(std::string("BorderedCR::") + what + ", bad leading dimension " + std::to_string(M.ldim()) + "\n");
}

// hxx:407-423: the new partition applies from the next factorization
template <typename T>
void BorderedCR<T>::set_num_parallel_block(integer npb) {
  max_pb_ = std::max<integer>(1, npb);
  if (gpu_) gpu_->want_P = max_pb_;
  factorized_ = false;
}

// cc:139-158: P shrinks until 2P <= N; 1 means pure cyclic reduction
template <typename T>
integer BorderedCR<T>::used_parallel_block() const {
  integer P = max_pb_, red = 2 * max_pb_;
  while (red > d_N_) { red -= 2; --P; }
  return P <= 1 ? 1 : P;
}

// cc:42-170 (the device buffers follow on first use)
template <typename T>
void BorderedCR<T>::allocate(integer N, integer n, integer qr, integer qx, integer nr, integer nx) {
  if (N < 1 || n < 1 || qr < 0 || qx < 0 || nr < 0 || nx < 0)
    fail("BorderedCR::allocate, bad dimensions nblock = " + std::to_string(N) + ", n = " + std::to_string(n) +
         ", qr = " + std::to_string(qr) + ", qx = " + std::to_string(qx) + ", nr = " + std::to_string(nr) +
         ", nx = " + std::to_string(nx) + "\n");
  if (N == d_N_ && n == d_n_ && qr == d_qr_ && qx == d_qx_ && nr == d_nr_ && nx == d_nx_ && !D_.empty())
    return;   // cc:55-61
  d_N_ = N; d_n_ = n; d_qr_ = qr; d_qx_ = qx; d_nr_ = nr; d_nx_ = nx;
  try {
    D_.assign(size_t(N) * n * n, T(0));
    E_.assign(size_t(N) * n * n, T(0));
    B_.assign(size_t(N) * n * nx, T(0));
    C_.assign(size_t(N + 1) * nr * n, T(0));
    Cq_.assign(size_t(nr) * qx, T(0));
    F_.assign(size_t(nr) * nx, T(0));
    H_.assign(size_t(n + qr) * Nc(), T(0));
  } catch (std::exception const& e) {
    last_error_ = e.what();
    throw;
  }
  device_bound_ = false;
  factorized_ = false;
}

// cc:181-198
template <typename T>
void BorderedCR<T>::dup(BorderedCR const& M) {
  set_num_parallel_block(M.max_pb_);
  allocate(M.d_N_, M.d_n_, M.d_qr_, M.d_qx_, M.d_nr_, M.d_nx_);
  D_ = M.D_; E_ = M.E_; B_ = M.B_; C_ = M.C_; Cq_ = M.Cq_; F_ = M.F_; H_ = M.H_;
  device_bound_ = false;
  factorized_ = false;
}

template <typename T>
void BorderedCR<T>::select(std::string_view s) {
  if (s == "LU") select_LU();
  else if (s == "QR") select_QR();
  else if (s == "QRP") select_QRP();
  else throw std::invalid_argument("BorderedCR::select( choice='" + std::string(s) +
                                   "' )\nunknown/unsupported factorization type ('LU','QR','QRP')");
}

template <typename T>
void BorderedCR<T>::select_last(std::string_view s) {
  static char const* names[] = {"LU", "LUPQ", "QR", "QRP", "SVD", "LSS", "LSY", "PINV"};
  for (int i = 0; i < 8; ++i)
    if (s == names[i]) { last_ = BORDERED_LAST_Choice(i); return; }
  throw std::invalid_argument("BorderedCR::select_last( choice='" + std::string(s) +
                              "' )\nunknown/unsupported factorization type ('LU','LUPQ','QR','QRP','SVD','LSS','LSY','PINV')");
}

template <typename T>
void BorderedCR<T>::select_last2(std::string_view s) {
  static char const* names[] = {"NONE", "SVD", "LSS", "LSY", "PINV"};
  for (int i = 0; i < 5; ++i)
    if (s == names[i]) { last2_ = BORDERED_LAST_Choice2(i); return; }
  throw std::invalid_argument("BorderedCR::select_last2( choice='" + std::string(s) +
                              "' )\nunknown/unsupported factorization type ('NONE','SVD','LSS','LSY','PINV')");
}

// hxx:548-590
template <typename T>
std::string BorderedCR<T>::choice_to_string(BORDERED_Choice c) {
  static char const* s[] = {"CyclicReduction+LU", "CyclicReduction+QR", "CyclicReduction+QRP"};
  return s[int(c)];
}
template <typename T>
std::string BorderedCR<T>::choice_to_string(BORDERED_LAST_Choice c) {
  static char const* s[] = {"LastBlock LU", "LastBlock LUPQ", "LastBlock QR", "LastBlock QRP",
                            "LastBlock SVD", "LastBlock LSS", "LastBlock LSY", "LastBlock PINV"};
  return s[int(c)];
}
template <typename T>
std::string BorderedCR<T>::choice_to_string(BORDERED_LAST_Choice2 c) {
  static char const* s[] = {"LastBlock NONE", "LastBlock SVD", "LastBlock LSS", "LastBlock LSY", "LastBlock PINV"};
  return s[int(c)];
}

template <typename T>
std::string BorderedCR<T>::info_algo() const {
  return choice_to_string(sel_) + " and " + choice_to_string(last_) + " and " + choice_to_string(last2_);
}

// cc:208-235; the reference prints the pool's thread count first: here the
// number of block ranges, which is what the threads compute in the reference
template <typename T>
std::string BorderedCR<T>::info(std::string_view ind) const {
  std::ostringstream s;
  std::string const I(ind);
  s << std::boolalpha << I << "rows   = " << nrows() << "\n" << I << "cols   = " << ncols() << "\n"
    << I << "nblock = " << d_N_ << "\n" << I << "n      = " << d_n_ << "\n" << I << "qr     = " << d_qr_ << "\n"
    << I << "qx     = " << d_qx_ << "\n" << I << "nr     = " << d_nr_ << "\n" << I << "nx     = " << d_nx_ << "\n"
    << I << "treads = " << max_pb_ << "/nblk:" << used_parallel_block() << "/factorize:" << fac_thread_
    << "/solve:" << sol_thread_ << "\n"
    << I << "intern = " << choice_to_string(sel_) << "\n" << I << "last   = " << choice_to_string(last_) << "\n"
    << I << "last2  = " << choice_to_string(last2_) << "\n" << I << "device = MI355X (gfx950)\n";
  return s.str();
}

template <typename T> void BorderedCR<T>::zero_D() { std::fill(D_.begin(), D_.end(), T(0)); }
template <typename T> void BorderedCR<T>::zero_E() { std::fill(E_.begin(), E_.end(), T(0)); }
template <typename T> void BorderedCR<T>::zero_B() { std::fill(B_.begin(), B_.end(), T(0)); }
template <typename T> void BorderedCR<T>::zero_F() { std::fill(F_.begin(), F_.end(), T(0)); }
template <typename T> void BorderedCR<T>::zero_H() { std::fill(H_.begin(), H_.end(), T(0)); }
template <typename T> void BorderedCR<T>::zero_C() { std::fill(C_.begin(), C_.end(), T(0)); }
template <typename T> void BorderedCR<T>::zero_Cq() { std::fill(Cq_.begin(), Cq_.end(), T(0)); }

// ---- loaders (cc:237-793) ------------------------------------------------------

template <typename T>
void BorderedCR<T>::load_B(integer k, T const B[], integer ld) {
  check_ld("load_B", k, ld, d_n_);
  gecopy(d_n_, d_nx_, B, ld, &B_[size_t(k) * d_n_ * d_nx_], d_n_);
}
template <typename T>
void BorderedCR<T>::load_B(integer k, MatW const& B) {
  check_dims("load_B", B, d_n_, d_nx_);
  load_B(k, B.data(), B.ldim());
}
template <typename T>
void BorderedCR<T>::add_to_B(integer k, T const B[], integer ld) {
  check_ld("add_to_B", k, ld, d_n_);
  geadd(d_n_, d_nx_, B, ld, &B_[size_t(k) * d_n_ * d_nx_], d_n_);
}
template <typename T>
void BorderedCR<T>::add_to_B(integer k, MatW const& B) {
  check_dims("add_to_B", B, d_n_, d_nx_);
  add_to_B(k, B.data(), B.ldim());
}
template <typename T>
void BorderedCR<T>::load_C(integer k, T const C[], integer ld) {
  check_ld("load_C", k, ld, d_nr_);
  gecopy(d_nr_, d_n_, C, ld, &C_[size_t(k) * d_nr_ * d_n_], d_nr_);
}
template <typename T>
void BorderedCR<T>::load_C(integer k, MatW const& C) {
  check_dims("load_C", C, d_nr_, d_n_);
  load_C(k, C.data(), C.ldim());
}
template <typename T>
void BorderedCR<T>::add_to_C(integer k, T const C[], integer ld) {
  check_ld("add_to_C", k, ld, d_nr_);
  geadd(d_nr_, d_n_, C, ld, &C_[size_t(k) * d_nr_ * d_n_], d_nr_);
}
template <typename T>
void BorderedCR<T>::add_to_C(integer k, MatW const& C) {
  check_dims("add_to_C", C, d_nr_, d_n_);
  add_to_C(k, C.data(), C.ldim());
}
// C_k and C_{k+1} side by side (nr x 2n)
template <typename T>
void BorderedCR<T>::add_to_C2(integer k, T const C[], integer ld) {
  check_ld("add_to_C2", k, ld, d_nr_);
  geadd(d_nr_, 2 * d_n_, C, ld, &C_[size_t(k) * d_nr_ * d_n_], d_nr_);
}
template <typename T>
void BorderedCR<T>::add_to_C2(integer k, MatW const& C) {
  check_dims("add_to_C2", C, d_nr_, 2 * d_n_);
  add_to_C2(k, C.data(), C.ldim());
}
// [C_k | C_{k+1} | F] (nr x (2n + nx))
template <typename T>
void BorderedCR<T>::add_to_C2F(integer k, T const C2F[], integer ld) {
  check_ld("add_to_C2F", k, ld, d_nr_);
  geadd(d_nr_, 2 * d_n_, C2F, ld, &C_[size_t(k) * d_nr_ * d_n_], d_nr_);
  geadd(d_nr_, d_nx_, C2F + size_t(ld) * 2 * d_n_, ld, F_.data(), d_nr_);
}
template <typename T>
void BorderedCR<T>::add_to_C2F(integer k, MatW const& C2F) {
  check_dims("add_to_C2F", C2F, d_nr_, 2 * d_n_ + d_nx_);
  add_to_C2F(k, C2F.data(), C2F.ldim());
}
template <typename T>
void BorderedCR<T>::load_D(integer k, T const D[], integer ld) {
  check_ld("load_D", k, ld, d_n_);
  gecopy(d_n_, d_n_, D, ld, &D_[size_t(k) * d_n_ * d_n_], d_n_);
}
template <typename T>
void BorderedCR<T>::load_D(integer k, MatW const& D) {
  check_dims("load_D", D, d_n_, d_n_);
  load_D(k, D.data(), D.ldim());
}
template <typename T>
void BorderedCR<T>::load_E(integer k, T const E[], integer ld) {
  check_ld("load_E", k, ld, d_n_);
  gecopy(d_n_, d_n_, E, ld, &E_[size_t(k) * d_n_ * d_n_], d_n_);
}
template <typename T>
void BorderedCR<T>::load_E(integer k, MatW const& E) {
  check_dims("load_E", E, d_n_, d_n_);
  load_E(k, E.data(), E.ldim());
}
template <typename T>
void BorderedCR<T>::load_DE(integer k, T const DE[], integer ld) {
  load_D(k, DE, ld);
  load_E(k, DE + size_t(d_n_) * ld, ld);
}
template <typename T>
void BorderedCR<T>::load_DEB(integer k, T const DEB[], integer ld) {
  load_D(k, DEB, ld);
  load_E(k, DEB + size_t(d_n_) * ld, ld);
  load_B(k, DEB + size_t(2 * d_n_) * ld, ld);
}
template <typename T>
void BorderedCR<T>::load_F(T const F[], integer ld) {
  check_ld("load_F", 0, ld, d_nr_);
  gecopy(d_nr_, d_nx_, F, ld, F_.data(), d_nr_);
}
template <typename T>
void BorderedCR<T>::load_F(MatW const& F) {
  check_dims("load_F", F, d_nr_, d_nx_);
  load_F(F.data(), F.ldim());
}
template <typename T>
void BorderedCR<T>::add_to_F(T const F[], integer ld) {
  check_ld("add_to_F", 0, ld, d_nr_);
  geadd(d_nr_, d_nx_, F, ld, F_.data(), d_nr_);
}
template <typename T>
void BorderedCR<T>::add_to_F(MatW const& F) {
  check_dims("add_to_F", F, d_nr_, d_nx_);
  add_to_F(F.data(), F.ldim());
}
template <typename T>
void BorderedCR<T>::load_Cq(T const Cq[], integer ld) {
  check_ld("load_Cq", 0, ld, d_nr_);
  gecopy(d_nr_, d_qx_, Cq, ld, Cq_.data(), d_nr_);
}
template <typename T>
void BorderedCR<T>::load_Cq(MatW const& Cq) {
  check_dims("load_Cq", Cq, d_nr_, d_qx_);
  load_Cq(Cq.data(), Cq.ldim());
}
template <typename T>
void BorderedCR<T>::load_CqF(T const CqF[], integer ld) {
  check_ld("load_CqF", 0, ld, d_nr_);
  gecopy(d_nr_, d_qx_, CqF, ld, Cq_.data(), d_nr_);
  gecopy(d_nr_, d_nx_, CqF + size_t(d_qx_) * ld, ld, F_.data(), d_nr_);
}
// cc:289-305: H0Nqp = [H0 | HN | Hq | Hp], (n+qr) x Nc
template <typename T>
void BorderedCR<T>::load_bottom(T const H0[], integer ld0, T const HN[], integer ldN, T const Hq[], integer ldQ,
                                T const Hp[], integer ldP) {
  integer const m = d_n_ + d_qr_;
  check_ld("load_bottom(H0)", 0, ld0, m);
  check_ld("load_bottom(HN)", 0, ldN, m);
  if (d_qx_ > 0) check_ld("load_bottom(Hq)", 0, ldQ, m);
  if (d_nx_ > 0) check_ld("load_bottom(Hp)", 0, ldP, m);
  T* H = H_.data();
  gecopy(m, d_n_, H0, ld0, H, m); H += size_t(m) * d_n_;
  gecopy(m, d_n_, HN, ldN, H, m); H += size_t(m) * d_n_;
  gecopy(m, d_qx_, Hq, ldQ, H, m); H += size_t(m) * d_qx_;
  gecopy(m, d_nx_, Hp, ldP, H, m);
}
template <typename T>
void BorderedCR<T>::load_bottom(MatW const& H0, MatW const& HN, MatW const& Hq, MatW const& Hp) {
  integer const m = d_n_ + d_qr_;
  check_dims("load_bottom(H0)", H0, m, d_n_);
  check_dims("load_bottom(HN)", HN, m, d_n_);
  check_dims("load_bottom(Hq)", Hq, m, d_qx_);
  check_dims("load_bottom(Hp)", Hp, m, d_nx_);
  load_bottom(H0.data(), H0.ldim(), HN.data(), HN.ldim(), Hq.data(), Hq.ldim(), Hp.data(), Hp.ldim());
}
template <typename T>
void BorderedCR<T>::load_bottom(T const H0Nqp[], integer ldH) {
  check_ld("load_bottom", 0, ldH, d_n_ + d_qr_);
  gecopy(d_n_ + d_qr_, Nc(), H0Nqp, ldH, H_.data(), d_n_ + d_qr_);
}
template <typename T>
void BorderedCR<T>::load_bottom(MatW const& H) {
  check_dims("load_bottom", H, d_n_ + d_qr_, Nc());
  load_bottom(H.data(), H.ldim());
}
// cc:380-395
template <typename T>
void BorderedCR<T>::load_bottom2(T const C0[], integer ld0, T const CN[], integer ldN, T const Cq[], integer ldCq,
                                 T const F[], integer ldF) {
  check_ld("load_bottom2(C0)", 0, ld0, d_nr_);
  check_ld("load_bottom2(CN)", 0, ldN, d_nr_);
  if (d_qx_ > 0) check_ld("load_bottom2(Cq)", 0, ldCq, d_nr_);
  if (d_nx_ > 0) check_ld("load_bottom2(F)", 0, ldF, d_nr_);
  gecopy(d_nr_, d_n_, C0, ld0, C_.data(), d_nr_);
  gecopy(d_nr_, d_n_, CN, ldN, &C_[size_t(d_N_) * d_nr_ * d_n_], d_nr_);
  gecopy(d_nr_, d_qx_, Cq, ldCq, Cq_.data(), d_nr_);
  gecopy(d_nr_, d_nx_, F, ldF, F_.data(), d_nr_);
}
template <typename T>
void BorderedCR<T>::load_bottom2(MatW const& C0, MatW const& CN, MatW const& Cq, MatW const& F) {
  check_dims("load_bottom2(C0)", C0, d_nr_, d_n_);
  check_dims("load_bottom2(CN)", CN, d_nr_, d_n_);
  check_dims("load_bottom2(Cq)", Cq, d_nr_, d_qx_);
  check_dims("load_bottom2(F)", F, d_nr_, d_nx_);
  load_bottom2(C0.data(), C0.ldim(), CN.data(), CN.ldim(), Cq.data(), Cq.ldim(), F.data(), F.ldim());
}

template <typename T>
void BorderedCR<T>::load_bottom2(MatW const& H) {
  check_dims("load_bottom2", H, d_nr_, Nc());
  T const* p = H.data();
  const size_t ld = size_t(H.ldim());
  load_C(0, p, H.ldim());
  load_C(d_N_, p + ld * d_n_, H.ldim());
  load_Cq(p + ld * 2 * d_n_, H.ldim());
  load_F(p + ld * (2 * d_n_ + d_qx_), H.ldim());
}

// ---- factorize / solve ----------------------------------------------------------

namespace {
template <typename E>
void configure(E& e, int sel, int last, bool last2) {
  e.set_method(sel);
  e.set_last_method(last);
  e.set_last2(last2);
}
}  // namespace

template <typename T>
bool BorderedCR<T>::factorize() {
  factorized_ = false;
  try {
    if (debug_) check_matrix();
    if (d_N_ < 1) throw std::runtime_error("BorderedCR::factorize, not allocated");
    check_supported();
    Gpu& g = gpu();
    int32_t st = 0;
    device_bound_ = false;   // the host blocks become the operator
    if (mixed_) {
      if constexpr (std::is_same_v<T, double>) {
        auto& e = g.mx().engine();
        configure(e, int(sel_), int(last_), last2_ != BORDERED_LAST_Choice2::NONE);
        e.upload(D_.data(), E_.data(), B_.data(), C_.data(), H_.data(), Cq_.data(), F_.data());
        e.factorize_async();
        st = e.status_sync();
      }
    } else {
      auto& e = g.fp();
      configure(e, int(sel_), int(last_), last2_ != BORDERED_LAST_Choice2::NONE);
      e.upload(D_.data(), E_.data(), B_.data(), C_.data(), H_.data(), Cq_.data(), F_.data());
      e.factorize_async();
      st = e.status_sync();
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
bool BorderedCR<T>::solve(T x[]) const {
  return solve(1, x, std::max(nrows(), ncols()));
}

template <typename T>
bool BorderedCR<T>::solve(integer nrhs, T X[], integer ld) const {
  if (!factorized_) {
    last_error_ = "BorderedCR::solve, matrix not factorized";
    return false;
  }
  try {
    const integer L = std::max(nrows(), ncols());
    if (ld < L) throw std::invalid_argument("BorderedCR::solve, ldRhs < max(nrows,ncols)");
    if (nrhs < 0) throw std::invalid_argument("BorderedCR::solve, nrhs < 0");
    if (nrhs == 0) return true;
    Gpu& g = gpu();
    g.xbuf.resize(size_t(L) * nrhs);
    hipStream_t st = static_cast<hipStream_t>(stream());
    bcr::hip_check(hipMemcpy2DAsync(g.xbuf.p, size_t(L) * sizeof(T), X, size_t(ld) * sizeof(T), size_t(L) * sizeof(T),
                                    nrhs, hipMemcpyHostToDevice, st), "H2D x");
    solve_device(g.xbuf.p, nrhs, L, false);
    bcr::hip_check(hipMemcpy2DAsync(X, size_t(ld) * sizeof(T), g.xbuf.p, size_t(L) * sizeof(T), size_t(L) * sizeof(T),
                                    nrhs, hipMemcpyDeviceToHost, st), "D2H x");
    bcr::hip_check(hipStreamSynchronize(st), "solve sync");
    if (mixed_ && !bcr::refine_ok(g.report.status)) {
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
    mixed_ = true;
  } else {
    throw std::invalid_argument("BorderedCR::select_precision( mode='" + std::string(mode) +
                                "' ) unknown mode ('FP64','MIXED')");
  }
  if (!(tol > 0) || max_iter < 0) throw std::invalid_argument("BorderedCR::select_precision: tol > 0, max_iter >= 0");
  ir_tol_ = tol;
  ir_max_ = max_iter;
  factorized_ = false;
}

template <typename T>
RefineInfo BorderedCR<T>::refine_report() const {
  RefineInfo r;
  if (!gpu_) return r;
  r.iterations = gpu_->report.iterations;
  r.rel_residual = gpu_->report.rel_residual;
  r.first_residual = gpu_->report.first_residual;
  r.status = gpu_->report.status;
  r.backward_error = gpu_->report.backward_error;
  return r;
}

template <typename T>
void* BorderedCR<T>::stream() const {
  return static_cast<void*>(gpu().stream.get());
}

template <typename T>
void BorderedCR<T>::synchronize() const {
  bcr::hip_check(hipStreamSynchronize(gpu().stream.get()), "sync");
}

template <typename T>
void BorderedCR<T>::bind_device(T const* D, T const* E, T const* B, T const* C, T const* H, T const* Cq, T const* F) {
  if (d_N_ < 1) fail("BorderedCR::bind_device, not allocated");
  Gpu& g = gpu();
  g.fp().bind_device(D, E, B, C, H, Cq, F);
  if constexpr (std::is_same_v<T, double>)
    if (mixed_ || g.has_mix()) g.mx().engine().bind_device(D, E, B, C, H, Cq, F);
  device_bound_ = true;
  factorized_ = false;
}

template <typename T>
void BorderedCR<T>::check_supported() const {
  // the per-pair LDS tables of the sweep and generic kernels hold 64 pivots
  // and 16 border rows
  if (d_n_ > 64 || d_nr_ > 16)
    throw std::runtime_error("BorderedCR::factorize, block size n <= 64 and nr <= 16 are supported (n = " +
                             std::to_string(d_n_) + ", nr = " + std::to_string(d_nr_) + ")");
}

template <typename T>
int32_t BorderedCR<T>::factorize_device(bool sync) {
  if (!device_bound_) fail("BorderedCR::factorize_device: no device operator bound");
  try {
    check_supported();
  } catch (std::exception const& e) {
    fail(e.what());
  }
  Gpu& g = gpu();
  int32_t st = 0;
  if (mixed_) {
    if constexpr (std::is_same_v<T, double>) {
      auto& e = g.mx().engine();
      if (!e.has_operator()) fail("BorderedCR::factorize_device: no device operator bound");
      configure(e, int(sel_), int(last_), last2_ != BORDERED_LAST_Choice2::NONE);
      e.factorize_async();
      if (sync) st = e.status_sync();
    }
  } else {
    auto& e = g.fp();
    configure(e, int(sel_), int(last_), last2_ != BORDERED_LAST_Choice2::NONE);
    e.factorize_async();
    if (sync) st = e.status_sync();
  }
  st &= 3;   // bit 3 only records that the second last-block solver was used
  factorized_ = st == 0;
  return st;
}

namespace {

// X columns cols[0..) <-> the packed block Y (ldy), device to device
void gather_cols(double* Y, int64_t ldy, double const* X, int64_t ldx, std::vector<int> const& cols, int64_t len,
                 hipStream_t st) {
  for (size_t q = 0; q < cols.size(); ++q)
    bcr::hip_check(bcr::launch_convert<double, double>(X + size_t(cols[q]) * ldx, Y + q * size_t(ldy), len, st),
                   "gather");
}
void scatter_cols(double* X, int64_t ldx, double const* Y, int64_t ldy, std::vector<int> const& cols, int64_t len,
                  hipStream_t st) {
  for (size_t q = 0; q < cols.size(); ++q)
    bcr::hip_check(bcr::launch_convert<double, double>(Y + q * size_t(ldy), X + size_t(cols[q]) * ldx, len, st),
                   "scatter");
}

}  // namespace

// in-place solve of device columns; in MIXED mode each column is refined on
// its own, and the columns whose refinement fails are solved again by the
// FP64 QRP path when the fallback is on
template <typename T>
void BorderedCR<T>::solve_device(T* X, integer nrhs, int64_t ld, bool sync) const {
  if (!factorized_) fail("BorderedCR::solve, matrix not factorized");
  const int64_t L = std::max<int64_t>(nrows(), ncols());
  if (nrhs < 0) fail("BorderedCR::solve, nrhs < 0");
  if (nrhs > 1 && ld < L) fail("BorderedCR::solve, ldRhs < max(nrows,ncols)");
  if (nrhs == 0) return;
  Gpu& g = gpu();
  if (mixed_) {
    if constexpr (std::is_same_v<T, double>) {
      bcr::MixedEngine& m = g.mx();
      g.report = m.solve(X, nrhs, ld < L ? L : ld, ir_tol_, ir_max_);
      g.col_status = m.column_status();
      std::vector<int> bad;
      for (int c = 0; c < nrhs; ++c)
        if (!bcr::refine_ok(g.col_status[c])) bad.push_back(c);
      if (!bad.empty() && fallback_) {
        // FP64 cyclic reduction with QRP elimination on the pristine operator
        auto& e = g.fp();
        if (!device_bound_) e.upload(D_.data(), E_.data(), B_.data(), C_.data(), H_.data(), Cq_.data(), F_.data());
        configure(e, int(BORDERED_Choice::QRP), int(last_), last2_ != BORDERED_LAST_Choice2::NONE);
        e.factorize_async();
        const int32_t fst = e.status_sync() & 3;
        if (fst != 0) {
          for (int c : bad) g.col_status[c] = bcr::kFactorFailed;
        } else {
          hipStream_t st = e.stream();
          const int k = int(bad.size());
          g.fbx.resize(size_t(L) * k);
          gather_cols(g.fbx.p, L, m.rhs(), m.rhs_ld(), bad, L, st);
          e.template solve_async<double>(g.fbx.p, k, L);
          // FP64 refinement on the FP64 QRP factors (they lose digits on
          // nearly singular pair columns; the sweeps recover them); a sweep
          // that does not help is undone
          g.fbb.resize(size_t(L) * k);
          gather_cols(g.fbb.p, L, m.rhs(), m.rhs_ld(), bad, L, st);
          double const* B = g.fbb.p;
          std::vector<double> rel = m.rel_residuals(g.fbx.p, k, L, B, L);
          const int64_t nr = nrows();
          for (int it = 0; it < ir_max_; ++it) {
            double worst = 0;
            for (double r : rel) worst = std::max(worst, r == r ? r : INFINITY);
            if (worst <= ir_tol_) break;
            double* R = m.residuals();
            if (L > nr)
              for (int c = 0; c < k; ++c)
                bcr::hip_check(bcr::launch_convert<double, double>(nullptr, R + size_t(c) * L + nr, L - nr, st), "pad");
            e.template solve_async<double>(R, k, L);
            for (int c = 0; c < k; ++c)
              bcr::hip_check(bcr::launch_axpy(1.0, R + size_t(c) * L, g.fbx.p + size_t(c) * L, ncols(), st), "x += d");
            std::vector<double> nrel = m.rel_residuals(g.fbx.p, k, L, B, L);
            double nworst = 0;
            for (double r : nrel) nworst = std::max(nworst, r == r ? r : INFINITY);
            if (!(nworst < 0.9 * worst)) {
              // undo the sweep: R still holds the corrections
              for (int c = 0; c < k; ++c)
                bcr::hip_check(bcr::launch_axpy(-1.0, R + size_t(c) * L, g.fbx.p + size_t(c) * L, ncols(), st),
                               "x -= d");
              break;
            }
            rel = nrel;
          }
          // per column: tol reached, or the residual floor (backward error),
          // or a failure; the report's residual becomes that of the columns
          // as returned
          rel = m.rel_residuals(g.fbx.p, k, L, B, L);
          g.report.rel_residual = 0;
          for (int q = 0; q < k; ++q) {
            int s = bcr::kFallbackFailed;
            if (rel[q] <= ir_tol_) {
              s = bcr::kFallbackOk;
            } else if (rel[q] == rel[q]) {
              const double om = m.backward_error(g.fbx.p + size_t(q) * L, B + size_t(q) * L, m.residuals() + size_t(q) * L);
              if (!(om <= g.report.backward_error)) g.report.backward_error = om;
              if (om <= bcr::MixedEngine::kFloorBerr) s = bcr::kFloor;
            }
            g.col_status[bad[q]] = s;
            g.report.rel_residual = std::max(g.report.rel_residual, rel[q]);
          }
          scatter_cols(X, ld < L ? L : ld, g.fbx.p, L, bad, L, st);
        }
        configure(e, int(sel_), int(last_), last2_ != BORDERED_LAST_Choice2::NONE);
        // the reported residual is the worst over the columns as returned
        double worst = 0;
        for (int c = 0; c < nrhs; ++c) {
          const bool refb = std::find(bad.begin(), bad.end(), c) != bad.end();
          if (!refb) worst = std::max(worst, m.column_residual()[c]);
        }
        g.report.rel_residual = std::max(worst, g.report.rel_residual);
        int s = bcr::kConverged;
        for (int c : g.col_status) s = bcr::refine_worse(s, c);
        g.report.status = s;
      }
      if (sync) bcr::hip_check(hipStreamSynchronize(g.stream.get()), "solve sync");
    }
    return;
  }
  g.fp().template solve_async<T>(X, nrhs, ld < L ? L : ld);
  if (sync) bcr::hip_check(hipStreamSynchronize(g.stream.get()), "solve sync");
}

template <typename T>
void BorderedCR<T>::residual_device(double const* x, double const* b, double* r, bool sync) const {
  Gpu& g = gpu();
  hipStream_t st = g.stream.get();
  // the operator lives where the last factorization put it: the mixed
  // engine after a MIXED factorize of host blocks
  const bool in_fp = g.eng && g.eng->has_operator();
  const bool in_mix = g.has_mix() && g.mx().engine().has_operator();
  if (in_mix && (mixed_ || !in_fp)) {
    g.mx().engine().residual_async(x, b, r);
  } else if (in_fp) {
    g.fp().residual_async(x, b, r);
  } else if (device_bound_ || !D_.empty()) {
    // nothing factorized yet: the operator as loaded or bound
    if (!device_bound_) g.fp().upload(D_.data(), E_.data(), B_.data(), C_.data(), H_.data(), Cq_.data(), F_.data());
    g.fp().residual_async(x, b, r);
  } else {
    fail("BorderedCR::residual_device: no operator loaded or bound");
  }
  if (sync) bcr::hip_check(hipStreamSynchronize(st), "residual sync");
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

// ---- operator and COO format (cc:2111-2534) ---------------------------------------

template <typename T>
void BorderedCR<T>::Mv(T const x[], T res[]) const {
  for (integer i = 0; i < nrows(); ++i) res[i] = T(0);
  add_Mv(T(1), x, res);
}

template <typename T>
void BorderedCR<T>::add_Mv(T alpha, T const x[], T res[]) const {
  const integer N = d_N_, n = d_n_, nr = d_nr_, nx = d_nx_, qx = d_qx_;
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
  integer const m = n + d_qr_;
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
  const integer N = d_N_, n = d_n_, nr = d_nr_, nx = d_nx_, qr = d_qr_, qx = d_qx_;
  return n * N * (2 * n + nx + nr) + nr * (n + qx + nx) + (n + qr) * (2 * n + qx + nx);
}

// per-block patterns (cc:2178-2331): rows and columns of the block in A
template <typename T>
integer BorderedCR<T>::pattern_D(integer k, integer I[], integer J[], integer offs) const {
  integer kk = 0;
  for (integer ij = 0; ij < d_n_ * d_n_; ++ij, ++kk) { I[kk] = k * d_n_ + ij % d_n_ + offs; J[kk] = k * d_n_ + ij / d_n_ + offs; }
  return kk;
}
template <typename T>
integer BorderedCR<T>::pattern_E(integer k, integer I[], integer J[], integer offs) const {
  integer kk = 0;
  for (integer ij = 0; ij < d_n_ * d_n_; ++ij, ++kk) {
    I[kk] = k * d_n_ + ij % d_n_ + offs;
    J[kk] = (k + 1) * d_n_ + ij / d_n_ + offs;
  }
  return kk;
}
template <typename T>
integer BorderedCR<T>::pattern_B(integer k, integer I[], integer J[], integer offs) const {
  integer kk = 0;
  const integer c0 = (d_N_ + 1) * d_n_ + d_qx_;
  for (integer ij = 0; ij < d_n_ * d_nx_; ++ij, ++kk) { I[kk] = k * d_n_ + ij % d_n_ + offs; J[kk] = c0 + ij / d_n_ + offs; }
  return kk;
}
template <typename T>
integer BorderedCR<T>::pattern_H(integer I[], integer J[], integer offs) const {
  const integer nqr = d_n_ + d_qr_;
  integer kk = 0;
  for (integer ij = 0; ij < nqr * Nc(); ++ij, ++kk) {
    integer j = ij / nqr;
    if (j >= d_n_) j += d_N_ * d_n_ - d_n_;
    I[kk] = d_N_ * d_n_ + ij % nqr + offs;
    J[kk] = j + offs;
  }
  return kk;
}
template <typename T>
integer BorderedCR<T>::pattern_C(integer k, integer I[], integer J[], integer offs) const {
  const integer i0 = (d_N_ + 1) * d_n_ + d_qr_;
  integer kk = 0;
  for (integer ij = 0; ij < d_nr_ * d_n_; ++ij, ++kk) { I[kk] = i0 + ij % d_nr_ + offs; J[kk] = k * d_n_ + ij / d_nr_ + offs; }
  return kk;
}
template <typename T>
integer BorderedCR<T>::pattern_Cq(integer I[], integer J[], integer offs) const {
  const integer i0 = (d_N_ + 1) * d_n_ + d_qr_;
  integer kk = 0;
  for (integer ij = 0; ij < d_nr_ * d_qx_; ++ij, ++kk) {
    I[kk] = i0 + ij % d_nr_ + offs;
    J[kk] = (d_N_ + 1) * d_n_ + ij / d_nr_ + offs;
  }
  return kk;
}
template <typename T>
integer BorderedCR<T>::pattern_F(integer I[], integer J[], integer offs) const {
  const integer i0 = (d_N_ + 1) * d_n_ + d_qr_;
  integer kk = 0;
  for (integer ij = 0; ij < d_nr_ * d_nx_; ++ij, ++kk) {
    I[kk] = i0 + ij % d_nr_ + offs;
    J[kk] = (d_N_ + 1) * d_n_ + d_qx_ + ij / d_nr_ + offs;
  }
  return kk;
}
template <typename T>
integer BorderedCR<T>::values_D(integer k, T V[]) const {
  std::copy_n(&D_[size_t(k) * d_n_ * d_n_], size_t(d_n_) * d_n_, V);
  return d_n_ * d_n_;
}
template <typename T>
integer BorderedCR<T>::values_E(integer k, T V[]) const {
  std::copy_n(&E_[size_t(k) * d_n_ * d_n_], size_t(d_n_) * d_n_, V);
  return d_n_ * d_n_;
}
template <typename T>
integer BorderedCR<T>::values_B(integer k, T V[]) const {
  std::copy_n(&B_[size_t(k) * d_n_ * d_nx_], size_t(d_n_) * d_nx_, V);
  return d_n_ * d_nx_;
}
template <typename T>
integer BorderedCR<T>::values_C(integer k, T V[]) const {
  std::copy_n(&C_[size_t(k) * d_nr_ * d_n_], size_t(d_nr_) * d_n_, V);
  return d_nr_ * d_n_;
}
template <typename T>
integer BorderedCR<T>::values_H(T V[]) const {
  std::copy(H_.begin(), H_.end(), V);
  return integer(H_.size());
}
template <typename T>
integer BorderedCR<T>::values_Cq(T V[]) const {
  std::copy(Cq_.begin(), Cq_.end(), V);
  return integer(Cq_.size());
}
template <typename T>
integer BorderedCR<T>::values_F(T V[]) const {
  std::copy(F_.begin(), F_.end(), V);
  return integer(F_.size());
}

// cc:2342-2378: D, E, B per block row, then H, then C_0..C_N, Cq, F
template <typename T>
void BorderedCR<T>::sparse_pattern(integer I[], integer J[], integer offs) const {
  integer kk = 0;
  for (integer k = 0; k < d_N_; ++k) {
    kk += pattern_D(k, I + kk, J + kk, offs);
    kk += pattern_E(k, I + kk, J + kk, offs);
    kk += pattern_B(k, I + kk, J + kk, offs);
  }
  kk += pattern_H(I + kk, J + kk, offs);
  for (integer k = 0; k <= d_N_; ++k) kk += pattern_C(k, I + kk, J + kk, offs);
  kk += pattern_Cq(I + kk, J + kk, offs);
  kk += pattern_F(I + kk, J + kk, offs);
  if (kk != sparse_nnz()) fail("BorderedCR::sparse_pattern( I, J, offs ), inserted values count mismatch\n");
}

template <typename T>
void BorderedCR<T>::sparse_values(T V[]) const {
  integer kk = 0;
  for (integer k = 0; k < d_N_; ++k) {
    kk += values_D(k, V + kk);
    kk += values_E(k, V + kk);
    kk += values_B(k, V + kk);
  }
  kk += values_H(V + kk);
  for (integer k = 0; k <= d_N_; ++k) kk += values_C(k, V + kk);
  kk += values_Cq(V + kk);
  kk += values_F(V + kk);
  if (kk != sparse_nnz()) fail("BorderedCR::sparse_values( V ), inserted values count mismatch\n");
}

// cc:2417-2482
template <typename T>
void BorderedCR<T>::sparse_load(T const V[], integer const Ir[], integer r_offs, integer const Jc[], integer c_offs,
                                integer nnz, bool do_error) {
  const integer N = d_N_, n = d_n_, nr = d_nr_, qr = d_qr_, qx = d_qx_, nx = d_nx_;
  const integer rH = n * N, rC = rH + n + qr, nrow = rC + nr;
  const integer cCq = rH + n, cF = cCq + qx, ncol = cF + nx;
  fill_zero();
  for (integer k = 0; k < nnz; ++k) {
    const integer i = Ir[k] - r_offs, j = Jc[k] - c_offs;
    const T v = V[k];
    bool ok = i >= 0 && j >= 0;
    const integer ib = ok ? i / n : 0, jb = ok ? j / n : 0;
    if (!ok) {
    } else if (i < rH) {
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
  factorized_ = false;
}

// cc:2493-2534: the matrix as a MATLAB sparse() script (1-based indices,
// 20 entries per line)
template <typename T>
void BorderedCR<T>::print_matlab_script(std::ostream& stream) const {
  const integer nnz = sparse_nnz();
  std::vector<integer> I(static_cast<size_t>(nnz)), J(static_cast<size_t>(nnz));
  std::vector<T> V(static_cast<size_t>(nnz));
  sparse_pattern(I.data(), J.data(), 1);
  sparse_values(V.data());
  auto list = [&](char const* name, auto const& a) {
    stream << name << " = [ " << num(a[0]);
    for (integer i = 1; i < nnz; ++i) {
      if (i % 20 == 0) stream << ", ...\n  " << num(a[i]);
      else stream << ",  " << num(a[i]);
    }
    stream << " ];\n\n";
  };
  if (nnz > 0) {
    list("I", I);
    list("J", J);
    list("V", V);
  }
  stream << "nr  = " << nrows() << ";\nnc  = " << ncols() << ";\n\nnnz = " << nnz
         << ";\n\nMAT = sparse( I, J, V, nr, nc );\n";
}

// cc:2538-2659: NaN / Inf scan of the loaded blocks, throws on the first hit
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
