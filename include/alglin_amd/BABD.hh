// alglin_amd::BABD<T> -- drop-in for the cyclic-reduction half of the
// reference's high-level wrapper alglin::BABD<T> (src/Alglin/BABD.hxx:56-367).
//
// The reference wrapper holds either a DiazLU or a BorderedCR backend and
// forwards factorize_system / solve_system / sparse_* to the active one.  Only
// the BorderedCR backend exists here (the Diaz solver is another algorithm,
// not part of this library): the three CYCLIC_REDUCTION_* choices behave as in
// the reference (the choice picks the LU / QR / QRP elimination of
// BorderedCR, BABD.hxx:105-120), while DIAZ is refused with
// std::invalid_argument and diaz() throws like the reference's accessor on a
// cyclic-reduction backend (BABD.hxx:165-170).
#pragma once

#include <stdexcept>
#include <string>

#include "alglin_amd/BorderedCR.hh"

namespace alglin_amd {

template <typename T>
class BABD {
 public:
  using real_type = T;
  using CRBackend = BorderedCR<T>;

  // BABD.hxx:66-71 (same values: the C interface's mat_fact codes)
  enum class BABD_Choice : integer {
    DIAZ = 1,
    CYCLIC_REDUCTION_LU = 2,
    CYCLIC_REDUCTION_QR = 3,
    CYCLIC_REDUCTION_QRP = 4
  };

  static std::string Choice_to_string(BABD_Choice c) {
    switch (c) {
      case BABD_Choice::DIAZ: return "Diaz";
      case BABD_Choice::CYCLIC_REDUCTION_LU: return "CyclicReduction+LU";
      case BABD_Choice::CYCLIC_REDUCTION_QR: return "CyclicReduction+QR";
      case BABD_Choice::CYCLIC_REDUCTION_QRP: return "CyclicReduction+QRP";
    }
    return "none";
  }

  explicit BABD(BABD_Choice choice = BABD_Choice::CYCLIC_REDUCTION_LU, Utils::ThreadPoolBase* TP = nullptr,
                integer num_parallel_block = 0)
      : cr_(TP, num_parallel_block) {
    selectSolver(choice);
  }
  BABD(BABD const&) = delete;
  BABD& operator=(BABD const&) = delete;

  BABD_Choice selected_solver() const { return choice_; }
  bool using_diaz() const { return false; }
  bool using_cyclic_reduction() const { return true; }

  CRBackend& cr() { return cr_; }
  CRBackend const& cr() const { return cr_; }
  // the reference asserts (throws) when the active backend is not DiazLU
  CRBackend& diaz() { throw std::runtime_error("BABD::diaz(), active backend is not DiazLU\n"); }

  // BABD.hxx:188-200: the choice re-selects the elimination of BorderedCR
  void selectSolver(BABD_Choice choice) {
    switch (choice) {
      case BABD_Choice::CYCLIC_REDUCTION_LU: cr_.select_LU(); break;
      case BABD_Choice::CYCLIC_REDUCTION_QR: cr_.select_QR(); break;
      case BABD_Choice::CYCLIC_REDUCTION_QRP: cr_.select_QRP(); break;
      default:
        throw std::invalid_argument("BABD::selectSolver, choice '" + Choice_to_string(choice) +
                                    "' is not available (cyclic reduction only)");
    }
    choice_ = choice;
  }

  std::string info() const { return cr_.info_algo(); }
  integer nrows() const { return cr_.nrows(); }
  integer ncols() const { return cr_.ncols(); }

  // BABD.hxx:271-313
  bool factorize_system() { return cr_.factorize(); }
  bool solve_system(real_type in_out[]) const { return cr_.solve(in_out); }
  bool solve_system(integer nrhs, real_type in_out[], integer ldIO) const { return cr_.solve(nrhs, in_out, ldIO); }
  bool factorize() { return factorize_system(); }
  bool solve(real_type in_out[]) const { return solve_system(in_out); }
  bool solve(integer nrhs, real_type in_out[], integer ldIO) const { return solve_system(nrhs, in_out, ldIO); }

  // BABD.hxx:336-364 (0-based pattern, like the reference's CR branch)
  integer sparse_nnz() const { return cr_.sparse_nnz(); }
  void sparse_pattern(integer I[], integer J[]) const { cr_.sparse_pattern(I, J, 0); }
  void sparse_values(real_type vals[]) const { cr_.sparse_values(vals); }

 private:
  BABD_Choice choice_ = BABD_Choice::CYCLIC_REDUCTION_LU;
  CRBackend cr_;
};

}  // namespace alglin_amd
