// Mixed-precision iterative refinement (see bcr_refine.hh).
#include "bcr_refine.hh"

#include <algorithm>
#include <cmath>
#include <numeric>

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

std::vector<double> MixedEngine::rel_residuals(double const* X, int nrhs, int64_t ld, double const* B, int64_t ldb) {
  norms_async(X, nrhs, ld, B, ldb);
  hnorm_.resize(size_t(2) * nrhs);
  hipStream_t st = eng_.stream();
  hip_check(hipMemcpyAsync(hnorm_.data(), norms_.p, hnorm_.size() * sizeof(double), hipMemcpyDeviceToHost, st),
            "norms");
  hip_check(hipStreamSynchronize(st), "norms sync");
  std::vector<double> rel(static_cast<size_t>(nrhs));
  for (int c = 0; c < nrhs; ++c) {
    const double nb = hnorm_[2 * c] > 0 ? hnorm_[2 * c] : 1.0;
    rel[c] = hnorm_[2 * c + 1] / nb;
  }
  return rel;
}

double MixedEngine::max_rel_residual(double const* X, int nrhs, int64_t ld, double const* B, int64_t ldb) {
  double worst = 0;
  for (double rel : rel_residuals(X, nrhs, ld, B, ldb))
    if (!(rel <= worst)) worst = rel;   // NaN propagates
  return worst;
}

double MixedEngine::backward_error(double const* x, double const* b, double const* r) {
  const int64_t nrows = eng_.dims().nrows();
  den_.resize(size_t(nrows) + 1);
  eng_.abs_bound_async(x, b, den_.p);
  eng_.max_ratio_async(r, den_.p, nrows, den_.p + nrows);
  double om = 0;
  hipStream_t st = eng_.stream();
  hip_check(hipMemcpyAsync(&om, den_.p + nrows, sizeof(double), hipMemcpyDeviceToHost, st), "omega");
  hip_check(hipStreamSynchronize(st), "omega sync");
  return om;
}

RefineReport MixedEngine::solve(double* X, int nrhs, int64_t ld, double tol, int max_iter, double stall) {
  RefineReport rep;
  col_status_.assign(size_t(std::max(nrhs, 0)), kConverged);
  col_rel_.assign(size_t(std::max(nrhs, 0)), 0.0);
  if (nrhs <= 0) return rep;
  Dims const& d = eng_.dims();
  const int64_t nrows = d.nrows(), L = rhs_ld();
  hipStream_t st = eng_.stream();
  b_.resize(size_t(L) * nrhs);
  r_.resize(size_t(L) * nrhs);
  norms_.resize(size_t(2) * nrhs);
  hnorm_.resize(size_t(2) * nrhs);
  hip_check(hipMemcpy2DAsync(b_.p, size_t(L) * sizeof(double), X, size_t(ld) * sizeof(double),
                             size_t(L) * sizeof(double), nrhs, hipMemcpyDeviceToDevice, st),
            "copy b");
  eng_.solve_async<double>(X, nrhs, ld);
  std::vector<int> open(static_cast<size_t>(nrhs));
  std::iota(open.begin(), open.end(), 0);
  std::vector<double> best(static_cast<size_t>(nrhs), INFINITY);
  std::vector<int> stagnant(static_cast<size_t>(nrhs), 0);
  for (int it = 0;; ++it) {
    // residuals of every column in one pass over the blocks (column c ->
    // slot c of r_; a closed column's x no longer changes, so computing it
    // costs flops, not bytes); the norms of the open ones, open[i] -> 2i
    const int m = int(open.size());
    eng_.residual_multi_async(X, nrhs, ld, b_.p, L, r_.p, L);
    for (int i = 0; i < m; ++i) {
      const int c = open[i];
      eng_.nrm2_async(b_.p + size_t(c) * L, nrows, norms_.p + 2 * i);
      eng_.nrm2_async(r_.p + size_t(c) * L, nrows, norms_.p + 2 * i + 1);
    }
    hip_check(hipMemcpyAsync(hnorm_.data(), norms_.p, size_t(2) * m * sizeof(double), hipMemcpyDeviceToHost, st),
              "norms");
    hip_check(hipStreamSynchronize(st), "norms sync");
    std::vector<int> keep;
    for (int i = 0; i < m; ++i) {
      const int c = open[i];
      const double nb = hnorm_[2 * i] > 0 ? hnorm_[2 * i] : 1.0;
      const double rel = hnorm_[2 * i + 1] / nb;
      if (it == 0 && !(rel <= rep.first_residual)) rep.first_residual = rel;
      col_rel_[c] = rel;
      rep.iterations = it;
      int done = -1;
      if (!(rel == rel)) {
        done = kStalled;   // NaN
      } else if (rel <= tol) {
        done = kConverged;
      } else if (it >= max_iter) {
        done = kMaxIter;
      } else if (rel > stall * best[c]) {
        // a sweep that does not shrink the best residual by `stall`: the
        // backward error tells whether FP64 can get any further (the floor),
        // and three such sweeps in a row end the column as stalled
        ++stagnant[c];
        const double om = backward_error(X + size_t(c) * ld, b_.p + size_t(c) * L, r_.p + size_t(c) * L);
        if (!(om <= rep.backward_error)) rep.backward_error = om;
        if (om <= kFloorBerr) done = kFloor;
        else if (stagnant[c] >= 3) done = kStalled;
      } else {
        stagnant[c] = 0;
      }
      best[c] = std::min(best[c], rel);
      if (done >= 0) col_status_[c] = done;
      else keep.push_back(i);
    }
    if (keep.empty()) break;
    // the residuals of the columns still open become the correction
    // right-hand sides, packed to the front of r_ (slot open[keep[k]] -> k;
    // the open columns ascend, so k <= open[keep[k]] and no slot is
    // overwritten before it is read); rows beyond nrows (when ncols > nrows)
    // are zero
    std::vector<int> next;
    for (int k = 0; k < int(keep.size()); ++k) {
      const int c = open[keep[k]];
      if (c != k)
        hip_check(launch_convert<double, double>(r_.p + size_t(c) * L, r_.p + size_t(k) * L, nrows, st), "pack r");
      if (L > nrows)
        hip_check(launch_convert<double, double>(nullptr, r_.p + size_t(k) * L + nrows, L - nrows, st), "pad");
      next.push_back(c);
    }
    const int k = int(next.size());
    eng_.solve_async<double>(r_.p, k, L);
    for (int q = 0; q < k; ++q)
      hip_check(launch_axpy(1.0, r_.p + size_t(q) * L, X + size_t(next[q]) * ld, d.ncols(), st), "x += d");
    open.swap(next);
  }
  rep.status = kConverged;
  rep.rel_residual = 0;
  for (int c = 0; c < nrhs; ++c) {
    rep.status = refine_worse(rep.status, col_status_[c]);
    if (!(col_rel_[c] <= rep.rel_residual)) rep.rel_residual = col_rel_[c];
  }
  return rep;
}

}  // namespace bcr
