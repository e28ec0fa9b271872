// Mixed-precision iterative refinement (see bcr_refine.hh).
#include "bcr_refine.hh"

#include <algorithm>
#include <cmath>

#include "bcr_launch.hh"

namespace bcr {

void MixedEngine::allocate(Dims const& d) {
  eng_.allocate(d);
  norms_.resize(2);
}

int64_t MixedEngine::rhs_ld() const {
  Dims const& d = eng_.dims();
  return std::max(d.nrows(), d.ncols());
}

// norms_[2c] = ||B_c||, norms_[2c+1] = ||B_c - A X_c||; r_ holds the residuals
void MixedEngine::norms_async(double const* X, int nrhs, int64_t ld, double const* B, int64_t ldb) {
  const int64_t nrows = eng_.dims().nrows(), L = rhs_ld();
  r_.resize(size_t(L) * nrhs);
  norms_.resize(size_t(2) * nrhs);
  for (int c = 0; c < nrhs; ++c) {
    eng_.residual_async(X + size_t(c) * ld, B + size_t(c) * ldb, r_.p + size_t(c) * L);
    eng_.nrm2_async(B + size_t(c) * ldb, nrows, norms_.p + 2 * c);
    eng_.nrm2_async(r_.p + size_t(c) * L, nrows, norms_.p + 2 * c + 1);
  }
}

double MixedEngine::max_rel_residual(double const* X, int nrhs, int64_t ld, double const* B, int64_t ldb) {
  norms_async(X, nrhs, ld, B, ldb);
  hnorm_.resize(size_t(2) * nrhs);
  hipStream_t st = eng_.stream();
  hip_check(hipMemcpyAsync(hnorm_.data(), norms_.p, hnorm_.size() * sizeof(double), hipMemcpyDeviceToHost, st),
            "norms");
  hip_check(hipStreamSynchronize(st), "norms sync");
  double worst = 0;
  for (int c = 0; c < nrhs; ++c) {
    const double nb = hnorm_[2 * c] > 0 ? hnorm_[2 * c] : 1.0;
    const double rel = hnorm_[2 * c + 1] / nb;
    if (!(rel <= worst)) worst = rel;   // NaN propagates
  }
  return worst;
}

RefineReport MixedEngine::solve(double* X, int nrhs, int64_t ld, double tol, int max_iter, double stall) {
  RefineReport rep;
  if (nrhs <= 0) return rep;
  Dims const& d = eng_.dims();
  const int64_t nrows = d.nrows(), L = rhs_ld();
  hipStream_t st = eng_.stream();
  b_.resize(size_t(L) * nrhs);
  hip_check(hipMemcpy2DAsync(b_.p, size_t(L) * sizeof(double), X, size_t(ld) * sizeof(double),
                             size_t(L) * sizeof(double), nrhs, hipMemcpyDeviceToDevice, st),
            "copy b");
  eng_.solve_async<double>(X, nrhs, ld);
  double best = INFINITY;
  int stagnant = 0;
  for (int it = 0;; ++it) {
    const double rel = max_rel_residual(X, nrhs, ld, b_.p, L);
    if (it == 0) rep.first_residual = rel;
    rep.rel_residual = rel;
    rep.iterations = it;
    if (!(rel == rel)) { rep.status = 1; break; }   // NaN
    if (rel <= tol) { rep.status = 0; break; }
    if (it >= max_iter) { rep.status = 2; break; }
    // a sweep that does not shrink the best residual by `stall` counts as
    // stagnant; three in a row mean the FP64 noise floor (or divergence)
    if (rel > stall * best) {
      if (++stagnant >= 3) { rep.status = 1; break; }
    } else {
      stagnant = 0;
    }
    best = std::min(best, rel);
    // the residuals become the correction right-hand sides: rows beyond
    // nrows (when ncols > nrows) are zero
    if (L > nrows)
      for (int c = 0; c < nrhs; ++c)
        hip_check(launch_convert<double, double>(nullptr, r_.p + size_t(c) * L + nrows, L - nrows, st), "pad");
    eng_.solve_async<double>(r_.p, nrhs, L);
    for (int c = 0; c < nrhs; ++c)
      hip_check(launch_axpy(1.0, r_.p + size_t(c) * L, X + size_t(c) * ld, d.ncols(), st), "x += d");
  }
  return rep;
}

}  // namespace bcr
