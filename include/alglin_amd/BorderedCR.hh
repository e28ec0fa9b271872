// alglin_amd::BorderedCR<T> -- C++ drop-in for alglin::BorderedCR<T>
// (src/Alglin/BABD_BorderedCR.hxx:104-1176), the bordered cyclic-reduction
// solver of block-almost-bidiagonal (BABD) systems, computed on an MI355X.
//
// Same method names, argument meaning and error behaviour as the reference:
//  * the object owns its blocks; loaders copy caller blocks (column-major,
//    leading dimension) in the reference's internal layout (cc:237-793,
//    hxx:972-998); the MatW overloads check dimensions and throw;
//  * factorize() returns false and sets last_error() on a failed elimination
//    (cc:803-839); solve(x) works in place on a vector of max(nrows, ncols)
//    entries (cc:1489-1542); solve(nrhs, X, ld) on a column-major block
//    (cc:1546-1605);
//  * select*() with an unknown name, check_matrix() on a NaN/Inf entry,
//    t_solve() and the dimension-checked loaders throw (the reference throws
//    Utils::Runtime_Error, here std::runtime_error / std::invalid_argument);
//  * Mv / add_Mv / sparse_* / print_matlab_script follow cc:2111-2534.
//
// Deliberate differences:
//  * factorize() leaves the loaded blocks intact (the reference overwrites
//    D, E, B, C, F in place, cc:1099-1153): the FP64 operator stays available
//    for the residual of iterative refinement, and Mv after factorize still
//    applies A.
//  * The thread pool is not used (the GPU needs none).  num_parallel_block
//    sets the elimination order directly: P block ranges reduced on their own
//    and then the chain of their end rows (cc:139-158), the order the
//    reference's pool computes with P threads; P = 1 (the default, and what
//    the reference uses without a pool) is pure cyclic reduction.
//
// This header needs no HIP or ROCm headers: a C++ caller compiles it with
// any C++17 compiler and links libalglin_amd.so.
#pragma once

#include <cstddef>
#include <cstdint>
#include <memory>
#include <ostream>
#include <string>
#include <string_view>
#include <vector>

namespace Utils {
class ThreadPoolBase;   // the reference's pool type (accepted, not used)
}

namespace alglin_amd {

using integer = int32_t;

// Column-major matrix view with a leading dimension (the reference's
// lapack_wrapper::MatrixWrapper as BorderedCR's MatW overloads use it).
template <typename T>
class MatrixWrapper {
 public:
  MatrixWrapper() = default;
  MatrixWrapper(T* data, integer nr, integer nc, integer ld) { setup(data, nr, nc, ld); }
  void setup(T* data, integer nr, integer nc, integer ld) {
    data_ = data;
    nr_ = nr;
    nc_ = nc;
    ld_ = ld;
  }
  T* data() const { return data_; }
  integer nrows() const { return nr_; }
  integer ncols() const { return nc_; }
  integer ldim() const { return ld_; }
  T& operator()(integer i, integer j) const { return data_[i + std::size_t(j) * ld_]; }

 private:
  T* data_ = nullptr;
  integer nr_ = 0, nc_ = 0, ld_ = 0;
};

// Outcome of a refined (MIXED) solve, worst over the right-hand sides.
// status: 0 converged to tol; 1 stalled; 2 max_iter; 3 the FP64 fallback
// factorization failed; 4 solved by the FP64 QRP fallback to tol; 5 stopped
// at the FP64 residual floor (componentwise backward error <= 64 u); 6 the
// FP64 QRP fallback still missed tol.  0, 4 and 5 are successes.
struct RefineInfo {
  int iterations = 0;
  double rel_residual = 0;
  double first_residual = 0;
  int status = 0;
  double backward_error = -1;
};

template <typename T>
class BorderedCR {
 public:
  using real_type = T;
  using MatW = MatrixWrapper<T>;

  // hxx:110-136
  enum class BORDERED_Choice : integer { LU = 0, QR = 1, QRP = 2 };
  enum class BORDERED_LAST_Choice : integer { LU = 0, LUPQ = 1, QR = 2, QRP = 3, SVD = 4, LSS = 5, LSY = 6, PINV = 7 };
  enum class BORDERED_LAST_Choice2 : integer { NONE = 0, SVD = 1, LSS = 2, LSY = 3, PINV = 4 };

  // hxx:373-384: a null pool means one block range; otherwise
  // num_parallel_block ranges (0 counts as 1: the pool's thread count is not
  // available here)
  explicit BorderedCR(Utils::ThreadPoolBase* TP = nullptr, integer num_parallel_block = 0);
  ~BorderedCR();
  BorderedCR(BorderedCR const&) = delete;
  BorderedCR& operator=(BorderedCR const&) = delete;

  // hxx:391-423
  void set_factorize_use_thread(bool v) { fac_thread_ = v; }
  void set_solve_use_thread(bool v) { sol_thread_ = v; }
  bool factorize_use_thread() const { return fac_thread_; }
  bool solve_use_thread() const { return sol_thread_; }
  void set_num_parallel_block(integer num_parallel_block);
  integer num_parallel_block() const { return max_pb_; }
  integer used_parallel_block() const;

  void allocate(integer nblock, integer n, integer qr, integer qx, integer nr, integer nx);   // cc:42-170
  void dup(BorderedCR const& M);                                                               // cc:181-198
  std::string_view last_error() const { return last_error_; }

  // selection (hxx:458-540)
  void select(std::string_view s);
  void select_last(std::string_view s);
  void select_last2(std::string_view s);
  void select_LU() { sel_ = BORDERED_Choice::LU; }
  void select_QR() { sel_ = BORDERED_Choice::QR; }
  void select_QRP() { sel_ = BORDERED_Choice::QRP; }
  void select_last_LU() { last_ = BORDERED_LAST_Choice::LU; }
  void select_last_LUPQ() { last_ = BORDERED_LAST_Choice::LUPQ; }
  void select_last_QR() { last_ = BORDERED_LAST_Choice::QR; }
  void select_last_QRP() { last_ = BORDERED_LAST_Choice::QRP; }
  void select_last_SVD() { last_ = BORDERED_LAST_Choice::SVD; }
  void select_last_LSS() { last_ = BORDERED_LAST_Choice::LSS; }
  void select_last_LSY() { last_ = BORDERED_LAST_Choice::LSY; }
  void select_last_PINV() { last_ = BORDERED_LAST_Choice::PINV; }
  void select_last2_NONE() { last2_ = BORDERED_LAST_Choice2::NONE; }
  void select_last2_SVD() { last2_ = BORDERED_LAST_Choice2::SVD; }
  void select_last2_LSS() { last2_ = BORDERED_LAST_Choice2::LSS; }
  void select_last2_LSY() { last2_ = BORDERED_LAST_Choice2::LSY; }
  void select_last2_PINV() { last2_ = BORDERED_LAST_Choice2::PINV; }
  BORDERED_Choice selected() const { return sel_; }
  BORDERED_LAST_Choice last_selected() const { return last_; }
  BORDERED_LAST_Choice2 last_selected2() const { return last2_; }
  static std::string choice_to_string(BORDERED_Choice c);
  static std::string choice_to_string(BORDERED_LAST_Choice c);
  static std::string choice_to_string(BORDERED_LAST_Choice2 c);
  std::string info_algo() const;
  std::string info_algo_block() const { return choice_to_string(sel_); }
  std::string info_algo_last_block() const { return choice_to_string(last_); }
  std::string info_algo_last_block2() const { return choice_to_string(last2_); }
  std::string info(std::string_view indent = "") const;   // cc:208-235
  void info(std::ostream& stream) const { stream << info(); }
  void set_debug(bool d) { debug_ = d; }

  // dims (hxx:617-650)
  integer nrows() const { return d_n_ * (d_N_ + 1) + d_qr_ + d_nr_; }
  integer ncols() const { return d_n_ * (d_N_ + 1) + d_qx_ + d_nx_; }
  integer number_of_blocks() const { return d_N_; }
  integer block_size() const { return d_n_; }
  integer dim_qr() const { return d_qr_; }
  integer dim_qx() const { return d_qx_; }
  integer dim_nr() const { return d_nr_; }
  integer dim_nx() const { return d_nx_; }
  integer Nr() const { return 2 * d_n_ + d_nr_ + d_qr_; }
  integer Nc() const { return 2 * d_n_ + d_nx_ + d_qx_; }

  // zero / loaders (cc:237-793); pointer loaders require ld >= rows
  void zero_D();
  void zero_E();
  void zero_B();
  void zero_F();
  void zero_H();
  void zero_C();
  void zero_Cq();
  void fill_zero() { zero_B(); zero_C(); zero_Cq(); zero_D(); zero_E(); zero_F(); zero_H(); }
  void load_B(integer nbl, T const B[], integer ldB);
  void load_B(integer nbl, MatW const& B);
  void add_to_B(integer nbl, T const B[], integer ldB);
  void add_to_B(integer nbl, MatW const& B);
  void load_C(integer nbl, T const C[], integer ldC);
  void load_C(integer nbl, MatW const& C);
  void add_to_C(integer nbl, T const C[], integer ldC);
  void add_to_C(integer nbl, MatW const& C);
  void add_to_C2(integer nbl, T const C[], integer ldC);
  void add_to_C2(integer nbl, MatW const& C);
  void add_to_C2F(integer nbl, T const C2F[], integer ldC);
  void add_to_C2F(integer nbl, MatW const& C2F);
  void load_D(integer nbl, T const D[], integer ldD);
  void load_D(integer nbl, MatW const& D);
  void load_E(integer nbl, T const E[], integer ldE);
  void load_E(integer nbl, MatW const& E);
  void load_DE(integer nbl, T const DE[], integer ldDE);
  void load_DEB(integer nbl, T const DEB[], integer ldDEB);
  void load_F(T const F[], integer ldF);
  void load_F(MatW const& F);
  void add_to_F(T const F[], integer ldF);
  void add_to_F(MatW const& F);
  void load_Cq(T const Cq[], integer ldC);
  void load_Cq(MatW const& Cq);
  void load_CqF(T const CqF[], integer ldCF);
  void load_bottom(T const H0[], integer ld0, T const HN[], integer ldN, T const Hq[], integer ldQ, T const Hp[],
                   integer ldP);
  void load_bottom(MatW const& H0, MatW const& HN, MatW const& Hq, MatW const& Hp);
  void load_bottom(T const H0Nqp[], integer ldH);
  void load_bottom(MatW const& H);
  void load_bottom2(T const C0[], integer ld0, T const CN[], integer ldN, T const Cq[], integer ldCq, T const F[],
                    integer ldF);
  void load_bottom2(MatW const& C0, MatW const& CN, MatW const& Cq, MatW const& F);
  void load_bottom2(MatW const& H);   // nr x Nc: [C_0 | C_N | Cq | F] (cc:429-446)

  // element access (hxx:972-998) and block views (hxx:1002-1014)
  T& B(integer k, integer i, integer j) { return B_[std::size_t(k) * d_n_ * d_nx_ + i + std::size_t(j) * d_n_]; }
  T& C(integer k, integer i, integer j) { return C_[std::size_t(k) * d_nr_ * d_n_ + i + std::size_t(j) * d_nr_]; }
  T& D(integer k, integer i, integer j) { return D_[std::size_t(k) * d_n_ * d_n_ + i + std::size_t(j) * d_n_]; }
  T& E(integer k, integer i, integer j) { return E_[std::size_t(k) * d_n_ * d_n_ + i + std::size_t(j) * d_n_]; }
  T& F(integer i, integer j) { return F_[i + std::size_t(j) * d_nr_]; }
  T& Cq(integer i, integer j) { return Cq_[i + std::size_t(j) * d_nr_]; }
  T& H(integer i, integer j) { return H_[i + std::size_t(j) * (d_n_ + d_qr_)]; }
  T const& B(integer k, integer i, integer j) const { return B_[std::size_t(k) * d_n_ * d_nx_ + i + std::size_t(j) * d_n_]; }
  T const& C(integer k, integer i, integer j) const { return C_[std::size_t(k) * d_nr_ * d_n_ + i + std::size_t(j) * d_nr_]; }
  T const& D(integer k, integer i, integer j) const { return D_[std::size_t(k) * d_n_ * d_n_ + i + std::size_t(j) * d_n_]; }
  T const& E(integer k, integer i, integer j) const { return E_[std::size_t(k) * d_n_ * d_n_ + i + std::size_t(j) * d_n_]; }
  T const& F(integer i, integer j) const { return F_[i + std::size_t(j) * d_nr_]; }
  T const& Cq(integer i, integer j) const { return Cq_[i + std::size_t(j) * d_nr_]; }
  T const& H(integer i, integer j) const { return H_[i + std::size_t(j) * (d_n_ + d_qr_)]; }
  void B(integer k, MatW& w) { w.setup(&B_[std::size_t(k) * d_n_ * d_nx_], d_n_, d_nx_, d_n_); }
  void C(integer k, MatW& w) { w.setup(&C_[std::size_t(k) * d_nr_ * d_n_], d_nr_, d_n_, d_nr_); }
  void D(integer k, MatW& w) { w.setup(&D_[std::size_t(k) * d_n_ * d_n_], d_n_, d_n_, d_n_); }
  void E(integer k, MatW& w) { w.setup(&E_[std::size_t(k) * d_n_ * d_n_], d_n_, d_n_, d_n_); }
  void F(MatW& w) { w.setup(F_.data(), d_nr_, d_nx_, d_nr_); }
  void Cq(MatW& w) { w.setup(Cq_.data(), d_nr_, d_qx_, d_nr_); }
  void H(MatW& w) { w.setup(H_.data(), d_n_ + d_qr_, Nc(), d_n_ + d_qr_); }

  // raw host block arrays (reference internal layout)
  T* data_D() { return D_.data(); }
  T* data_E() { return E_.data(); }
  T* data_B() { return B_.data(); }
  T* data_C() { return C_.data(); }
  T* data_H() { return H_.data(); }
  T* data_Cq() { return Cq_.data(); }
  T* data_F() { return F_.data(); }

  // compute (GPU)
  bool factorize();                                     // cc:803-839
  bool solve(T x[]) const;                              // cc:1489-1542
  bool solve(integer nrhs, T rhs[], integer ldRhs) const;   // cc:1546-1605
  bool t_solve(T[]) const;                              // cc:2089-2101: throws
  bool t_solve(integer, T[], integer) const;

  // operator on the loaded blocks (cc:2111-2166), host side
  void Mv(T const x[], T res[]) const;
  void add_Mv(T alpha, T const x[], T res[]) const;

  // COO export / import (hxx:1079-1126, cc:2178-2482)
  integer sparse_nnz() const;
  void sparse_pattern(integer I[], integer J[], integer offs) const;
  void sparse_values(T V[]) const;
  void sparse_load(T const V[], integer const I[], integer r_offs, integer const J[], integer c_offs, integer nnz,
                   bool do_error = true);
  integer pattern_B(integer nbl, integer I[], integer J[], integer offs) const;
  integer values_B(integer nbl, T V[]) const;
  integer pattern_C(integer nbl, integer I[], integer J[], integer offs) const;
  integer values_C(integer nbl, T V[]) const;
  integer pattern_D(integer nbl, integer I[], integer J[], integer offs) const;
  integer values_D(integer nbl, T V[]) const;
  integer pattern_E(integer nbl, integer I[], integer J[], integer offs) const;
  integer values_E(integer nbl, T V[]) const;
  integer pattern_F(integer I[], integer J[], integer offs) const;
  integer values_F(T V[]) const;
  integer pattern_Cq(integer I[], integer J[], integer offs) const;
  integer values_Cq(T V[]) const;
  integer pattern_H(integer I[], integer J[], integer offs) const;
  integer values_H(T V[]) const;
  void print_matlab_script(std::ostream& stream) const;   // cc:2493-2534
  void check_matrix();                                    // cc:2538-2659

  // ---- extensions of the reference surface ----------------------------------
  // precision: "FP64" (default) or "MIXED" (double only): FP32 cyclic-reduction
  // factors + FP64 iterative refinement on the GPU to ||b - A x|| / ||b|| <= tol
  void select_precision(std::string_view mode, double tol = 1e-12, int max_iter = 30);
  bool mixed() const { return mixed_; }
  // MIXED: columns whose refinement stalls are solved again by FP64 cyclic
  // reduction with QRP elimination (on by default)
  void set_fallback(bool on) { fallback_ = on; }
  RefineInfo refine_report() const;
  // device-resident operation: caller-owned HIP device arrays (packed layout)
  // as the operator, device right-hand sides; stream() is the hipStream_t
  void bind_device(T const* D, T const* E, T const* B, T const* C, T const* H, T const* Cq, T const* F);
  bool device_bound() const { return device_bound_; }
  // 0 ok, bit 0 zero pivot, bit 1 singular last block (sync = false: 0, the
  // status is checked by the next synchronizing call)
  int32_t factorize_device(bool sync);
  void solve_device(T* X, integer nrhs, int64_t ld, bool sync) const;
  // r = b - A x in FP64 on the operator the object factors (double vectors)
  void residual_device(double const* x, double const* b, double* r, bool sync) const;
  void* stream() const;
  void synchronize() const;

  struct Gpu;   // device state (defined by the library)
  Gpu& gpu() const;

 private:
  void fail(std::string const& msg) const;
  void check_ld(char const* what, integer nbl, integer ld, integer rows) const;
  void check_dims(char const* what, MatW const& M, integer nr, integer nc) const;
  void check_supported() const;   // n <= 64, nr <= 16
  integer d_N_ = 0, d_n_ = 0, d_qr_ = 0, d_qx_ = 0, d_nr_ = 0, d_nx_ = 0;
  integer max_pb_ = 1;
  BORDERED_Choice sel_ = BORDERED_Choice::LU;
  BORDERED_LAST_Choice last_ = BORDERED_LAST_Choice::LU;
  BORDERED_LAST_Choice2 last2_ = BORDERED_LAST_Choice2::NONE;
  bool fac_thread_ = true, sol_thread_ = true, debug_ = false, factorized_ = false;
  bool mixed_ = false, fallback_ = true, device_bound_ = false;
  double ir_tol_ = 1e-12;
  int ir_max_ = 30;
  std::vector<T> D_, E_, B_, C_, Cq_, F_, H_;
  mutable std::unique_ptr<Gpu> gpu_;
  mutable std::string last_error_{"no error"};
};

}  // namespace alglin_amd
