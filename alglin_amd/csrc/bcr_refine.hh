// Mixed-precision solver: FP32 cyclic-reduction factors of the FP64 operator
// plus FP64 iterative refinement on the GPU.
//
//   x0 = M^{-1} b                      (M = FP32 factors, sweeps carried in FP64)
//   r_k = b - A x_k                    (FP64 block SpMV on the pristine A)
//   x_{k+1} = x_k + M^{-1} r_k
//
// until ||r_k||_2 / ||b||_2 <= tol, or three sweeps in a row fail to bring
// the best residual so far down by a factor `stall` (status = stalled: the
// FP64 noise floor of the operator is above tol), or max_iter.
// All vectors stay in HBM; the only host traffic per sweep is the two norms.
#pragma once

#include <vector>

#include "bcr_engine.hh"

namespace bcr {

struct RefineReport {
  int iterations = 0;        // correction sweeps after the first solve
  double rel_residual = 0;   // max over columns of ||b - A x|| / ||b|| at exit (FP64)
  double first_residual = 0; // the same after the first (FP32-factor) solve
  // 0 converged, 1 stalled, 2 max_iter, 3 factorization failed,
  // 4 refinement stalled and the FP64 QRP fallback produced x
  int status = 0;
};

class MixedEngine {
 public:
  MixedEngine() = default;
  Engine<double, float>& engine() { return eng_; }
  void allocate(Dims const& d);
  // X: device block (ld x nrhs), b on entry, solution on exit; all columns
  // are refined together (one multi-RHS sweep per iteration)
  RefineReport solve(double* X, int nrhs, int64_t ld, double tol, int max_iter, double stall = 0.5);
  // the right-hand sides of the last solve (ld = max(nrows, ncols))
  double const* rhs() const { return b_.p; }
  int64_t rhs_ld() const;
  // max over columns of ||B - A X|| / ||B|| (FP64, synchronizes the stream);
  // the residual columns stay in residuals() (ld rhs_ld())
  double max_rel_residual(double const* X, int nrhs, int64_t ld, double const* B, int64_t ldb);
  double* residuals() { return r_.p; }

 private:
  void norms_async(double const* X, int nrhs, int64_t ld, double const* B, int64_t ldb);
  Engine<double, float> eng_;
  DevBuf<double> b_, r_, norms_;
  std::vector<double> hnorm_;
};

}  // namespace bcr
