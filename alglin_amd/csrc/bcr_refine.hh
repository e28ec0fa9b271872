// Mixed-precision solver: FP32 cyclic-reduction factors of the FP64 operator
// plus FP64 iterative refinement on the GPU.
//
//   x0 = M^{-1} b                      (M = FP32 factors, sweeps carried in FP64)
//   r_k = b - A x_k                    (FP64 block SpMV on the pristine A)
//   x_{k+1} = x_k + M^{-1} r_k
//
// Each right-hand side is refined on its own: a column leaves the sweep set
// when ||r||_2 / ||b||_2 <= tol (converged) or when max_iter is reached.  A
// sweep that fails to bring the column's best residual down by a factor
// `stall` is judged by the componentwise backward error
// omega = max_i |r_i| / (|A| |x| + |b|)_i (Oettli-Prager, the measure LAPACK's
// dgerfs refines on): omega <= kFloorBerr means x is already an exact
// solution of a system within a few FP64 roundings of (A, b), so the residual
// cannot be pushed further in FP64 arithmetic ("FP64 floor", a success);
// three such sweeps in a row with omega above it end the column as stalled
// (a failure the caller may hand to the FP64 QRP path).  All vectors stay in HBM; per sweep only the residual norms of
// the open columns cross to the host.
#pragma once

#include <vector>

#include "bcr_engine.hh"

namespace bcr {

// per-column outcome of a refined solve
enum RefineStatus : int {
  kConverged = 0,      // ||r|| / ||b|| <= tol
  kStalled = 1,        // no progress and omega above the floor (or NaN)
  kMaxIter = 2,        // max_iter sweeps without reaching tol
  kFactorFailed = 3,   // the FP64 fallback factorization failed
  kFallbackOk = 4,     // the FP64 QRP fallback reached tol
  kFloor = 5,          // stopped at the FP64 residual floor: omega <= floor_berr
  kFallbackFailed = 6  // the FP64 QRP fallback ran and still missed tol
};

// true for the outcomes that return a solution to the caller
inline bool refine_ok(int status) { return status == kConverged || status == kFallbackOk || status == kFloor; }

// the outcome that describes a set of columns: any failure first, then the
// fallback, the floor, convergence
inline int refine_worse(int a, int b) {
  auto rank = [](int s) {
    switch (s) {
      case kFactorFailed: return 6;
      case kFallbackFailed: return 5;
      case kStalled: return 4;
      case kMaxIter: return 3;
      case kFallbackOk: return 2;
      case kFloor: return 1;
      default: return 0;
    }
  };
  return rank(a) >= rank(b) ? a : b;
}

struct RefineReport {
  int iterations = 0;        // most correction sweeps any column took
  double rel_residual = 0;   // max over columns of ||b - A x|| / ||b|| at exit (FP64)
  double first_residual = 0; // the same after the first (FP32-factor) solve
  int status = 0;            // worst RefineStatus over the columns
  double backward_error = -1; // max omega over the columns where it was evaluated (-1: none)
};

class MixedEngine {
 public:
  MixedEngine() = default;
  // on the caller's stream (not owned)
  explicit MixedEngine(hipStream_t st) : eng_(st) {}
  Engine<double, float>& engine() { return eng_; }
  void allocate(Dims const& d);
  // X: device block (ld x nrhs), b on entry, solution on exit
  RefineReport solve(double* X, int nrhs, int64_t ld, double tol, int max_iter, double stall = 0.5);
  // per-column outcome of the last solve
  std::vector<int> const& column_status() const { return col_status_; }
  std::vector<double> const& column_residual() const { return col_rel_; }
  // the right-hand sides of the last solve (ld = max(nrows, ncols))
  double const* rhs() const { return b_.p; }
  int64_t rhs_ld() const;
  // ||B_c - A X_c|| / ||B_c|| for every column (FP64, synchronizes the
  // stream); the residual columns stay in residuals() (ld rhs_ld())
  std::vector<double> rel_residuals(double const* X, int nrhs, int64_t ld, double const* B, int64_t ldb);
  double max_rel_residual(double const* X, int nrhs, int64_t ld, double const* B, int64_t ldb);
  // componentwise backward error of column x against b, with its residual
  // b - A x already in r (synchronizes the stream)
  double backward_error(double const* x, double const* b, double const* r);
  double* residuals() { return r_.p; }
  // omega at or below which a stalled column counts as solved to the FP64
  // floor: a small multiple of the FP64 unit roundoff
  static constexpr double kFloorBerr = 64 * 1.1102230246251565e-16;

 private:
  void norms_async(double const* X, int nrhs, int64_t ld, double const* B, int64_t ldb);
  Engine<double, float> eng_;
  DevBuf<double> b_, r_, norms_, den_;
  std::vector<double> hnorm_;
  std::vector<int> col_status_;
  std::vector<double> col_rel_;
};

}  // namespace bcr
