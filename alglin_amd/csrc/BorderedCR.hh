// alglin_amd::BorderedCR<T> -- the drop-in surface of alglin::BorderedCR<T>
// (src/Alglin/BABD_BorderedCR.hxx:104-1176) backed by the MI355X engine.
//
// Same method names, argument meaning and error behaviour as the reference:
// loaders copy caller blocks (column-major, leading dimension) into the
// object (cc:237-793); factorize() returns false and sets last_error() on a
// failed elimination (cc:803-839); solve() works in place on a vector of
// max(nrows, ncols) entries (cc:1489-1542); solve(nrhs, X, ld) on a
// column-major block (cc:1546-1605); Mv/add_Mv/sparse_* follow
// cc:2111-2482.  The compute of factorize/solve runs on the GPU; the thread
// pool arguments of the reference are accepted and ignored.
//
// Differences (documented, intentional):
//  * factorize() leaves the loaded blocks intact (the reference overwrites
//    D, E, B, C, F, cc:1099-1153); Mv after factorize therefore still
//    applies the original operator.
//  * the last block supports LU, LUPQ, QR and QRP on the GPU; SVD, LSS,
//    LSY and PINV are accepted by name and make factorize() report an error.
#pragma once

#include <memory>
#include <string>
#include <string_view>
#include <vector>

#include "bcr_engine.hh"
#include "bcr_refine.hh"

namespace alglin_amd {

using integer = int32_t;

template <typename T>
class BorderedCR {
 public:
  enum class Choice : integer { LU = 0, QR = 1, QRP = 2 };
  enum class LastChoice : integer { LU = 0, LUPQ = 1, QR = 2, QRP = 3, SVD = 4, LSS = 5, LSY = 6, PINV = 7 };
  enum class LastChoice2 : integer { NONE = 0, SVD = 1, LSS = 2, LSY = 3, PINV = 4 };

  explicit BorderedCR(void* thread_pool = nullptr, integer num_parallel_block = 0);
  ~BorderedCR();
  BorderedCR(BorderedCR const&) = delete;
  BorderedCR& operator=(BorderedCR const&) = delete;

  void allocate(integer nblock, integer n, integer qr, integer qx, integer nr, integer nx);
  void dup(BorderedCR const& M);
  std::string const& last_error() const { return last_error_; }

  // selection (hxx:458-540); strings throw std::invalid_argument like UTILS_ERROR
  void select(std::string_view s);
  void select_last(std::string_view s);
  void select_last2(std::string_view s);
  void select_LU() { sel_ = Choice::LU; }
  void select_QR() { sel_ = Choice::QR; }
  void select_QRP() { sel_ = Choice::QRP; }
  void select_last_LU() { last_ = LastChoice::LU; }
  void select_last_LUPQ() { last_ = LastChoice::LUPQ; }
  void select_last_QR() { last_ = LastChoice::QR; }
  void select_last_QRP() { last_ = LastChoice::QRP; }
  void select_last_SVD() { last_ = LastChoice::SVD; }
  void select_last_LSS() { last_ = LastChoice::LSS; }
  void select_last_LSY() { last_ = LastChoice::LSY; }
  void select_last_PINV() { last_ = LastChoice::PINV; }
  void select_last2_NONE() { last2_ = LastChoice2::NONE; }
  void select_last2_SVD() { last2_ = LastChoice2::SVD; }
  void select_last2_LSS() { last2_ = LastChoice2::LSS; }
  void select_last2_LSY() { last2_ = LastChoice2::LSY; }
  void select_last2_PINV() { last2_ = LastChoice2::PINV; }
  Choice selected() const { return sel_; }
  LastChoice last_selected() const { return last_; }
  std::string info_algo() const;
  std::string info(std::string_view indent = "") const;

  // thread knobs of the reference (accepted; the GPU path has no host pool)
  void set_factorize_use_thread(bool v) { fac_thread_ = v; }
  void set_solve_use_thread(bool v) { sol_thread_ = v; }
  void set_num_parallel_block(integer) {}
  void set_debug(bool d) { debug_ = d; }

  // dims (hxx:617-650)
  integer nrows() const { return integer(d_.nrows()); }
  integer ncols() const { return integer(d_.ncols()); }
  integer number_of_blocks() const { return d_.N; }
  integer block_size() const { return d_.n; }
  integer dim_qr() const { return d_.qr; }
  integer dim_qx() const { return d_.qx; }
  integer dim_nr() const { return d_.nr; }
  integer dim_nx() const { return d_.nx; }
  integer Nr() const { return d_.Nr(); }
  integer Nc() const { return d_.Nc(); }

  // zero / loaders (cc:237-793)
  void zero_D(); void zero_E(); void zero_B(); void zero_F(); void zero_H(); void zero_C(); void zero_Cq();
  void fill_zero() { zero_B(); zero_C(); zero_Cq(); zero_D(); zero_E(); zero_F(); zero_H(); }
  void load_B(integer nbl, T const B[], integer ldB);
  void add_to_B(integer nbl, T const B[], integer ldB);
  void load_C(integer nbl, T const C[], integer ldC);
  void add_to_C(integer nbl, T const C[], integer ldC);
  void add_to_C2(integer nbl, T const C[], integer ldC);
  void add_to_C2F(integer nbl, T const C2F[], integer ldC);
  void load_D(integer nbl, T const D[], integer ldD);
  void load_E(integer nbl, T const E[], integer ldE);
  void load_DE(integer nbl, T const DE[], integer ldDE);
  void load_DEB(integer nbl, T const DEB[], integer ldDEB);
  void load_F(T const F[], integer ldF);
  void add_to_F(T const F[], integer ldF);
  void load_Cq(T const Cq[], integer ldC);
  void load_CqF(T const CqF[], integer ldCF);
  void load_bottom(T const H0[], integer ld0, T const HN[], integer ldN, T const Hq[], integer ldQ, T const Hp[],
                   integer ldP);
  void load_bottom(T const H0Nqp[], integer ldH);
  void load_bottom2(T const C0[], integer ld0, T const CN[], integer ldN, T const Cq[], integer ldCq, T const F[],
                    integer ldF);

  // element access (hxx:972-998)
  T& B(integer k, integer i, integer j) { return B_[size_t(k) * d_.n * d_.nx + i + size_t(j) * d_.n]; }
  T& C(integer k, integer i, integer j) { return C_[size_t(k) * d_.nr * d_.n + i + size_t(j) * d_.nr]; }
  T& D(integer k, integer i, integer j) { return D_[size_t(k) * d_.n * d_.n + i + size_t(j) * d_.n]; }
  T& E(integer k, integer i, integer j) { return E_[size_t(k) * d_.n * d_.n + i + size_t(j) * d_.n]; }
  T& F(integer i, integer j) { return F_[i + size_t(j) * d_.nr]; }
  T& Cq(integer i, integer j) { return Cq_[i + size_t(j) * d_.nr]; }
  T& H(integer i, integer j) { return H_[i + size_t(j) * (d_.n + d_.qr)]; }

  // raw host block arrays (reference internal layout)
  T* data_D() { return D_.data(); }
  T* data_E() { return E_.data(); }
  T* data_B() { return B_.data(); }
  T* data_C() { return C_.data(); }
  T* data_H() { return H_.data(); }
  T* data_Cq() { return Cq_.data(); }
  T* data_F() { return F_.data(); }

  // compute (GPU)
  bool factorize();
  bool solve(T x[]);
  bool solve(integer nrhs, T rhs[], integer ldRhs);
  bool t_solve(T[]) const;                  // not defined in the reference either (cc:2089-2101): throws
  bool t_solve(integer, T[], integer) const;

  // operator on the loaded blocks (cc:2111-2166), host side
  void Mv(T const x[], T res[]) const;
  void add_Mv(T alpha, T const x[], T res[]) const;

  // COO export / import (cc:2178-2482)
  integer sparse_nnz() const;
  void sparse_pattern(integer I[], integer J[], integer offs) const;
  void sparse_values(T V[]) const;
  void sparse_load(T const V[], integer const I[], integer r_offs, integer const J[], integer c_offs, integer nnz,
                   bool do_error = true);
  void check_matrix();   // cc:2538-2659 NaN/Inf scan, throws

  bcr::Engine<T, T>& engine() { return *eng_; }

  // ---- precision (extension; the reference surface is the "FP64" default) --
  // "MIXED" (double only): FP32 cyclic-reduction factors + FP64 iterative
  // refinement on the GPU to ||b - A x|| / ||b|| <= tol.
  void select_precision(std::string_view mode, double tol = 1e-12, int max_iter = 30);
  // MIXED only: when refinement stalls, solve again with FP64 QRP elimination
  // (report status 4); off: the solve reports the failure instead
  void set_fallback(bool on) { fallback_ = on; }
  bool mixed() const { return mixed_; }
  bcr::RefineReport const& refine_report() const { return report_; }
  bcr::MixedEngine* mixed_engine() { return mix_.get(); }
  // device-resident entry points honouring the precision mode
  void bind_device(T const* D, T const* E, T const* B, T const* C, T const* H, T const* Cq, T const* F);
  int32_t factorize_device(bool sync);
  void solve_device(T* X, integer nrhs, int64_t ld, bool sync);

 private:
  void fail(std::string const& msg) const;
  bcr::Dims d_;
  Choice sel_ = Choice::LU;
  LastChoice last_ = LastChoice::LU;
  LastChoice2 last2_ = LastChoice2::NONE;
  bool fac_thread_ = true, sol_thread_ = true, debug_ = false, factorized_ = false;
  std::vector<T> D_, E_, B_, C_, Cq_, F_, H_;
  std::unique_ptr<bcr::Engine<T, T>> eng_;
  std::unique_ptr<bcr::MixedEngine> mix_;
  bool mixed_ = false;
  double ir_tol_ = 1e-12;
  int ir_max_ = 30;
  bool fallback_ = true, device_bound_ = false;
  void fallback_qrp(T* X, integer nrhs, int64_t ld);
  bcr::RefineReport report_;
  bcr::DevBuf<T> xbuf_;
  mutable std::string last_error_{"no error"};
};

}  // namespace alglin_amd
