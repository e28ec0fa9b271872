// Mixed-precision iterative refinement (see bcr_refine.hh).
#include "bcr_refine.hh"

#include <algorithm>
#include <cmath>

#include "bcr_launch.hh"

namespace bcr {

void MixedEngine::allocate(Dims const& d) {
  eng_.allocate(d);
  const size_t L = size_t(std::max(d.nrows(), d.ncols()));
  b_.resize(L);
  r_.resize(L);
  norms_.resize(4);
}

RefineReport MixedEngine::solve(double* x, double tol, int max_iter, double stall) {
  RefineReport rep;
  Dims const& d = eng_.dims();
  const int64_t nrows = d.nrows(), L = std::max(d.nrows(), d.ncols());
  hipStream_t st = eng_.stream();
  hip_check(hipMemcpyAsync(b_.p, x, size_t(L) * sizeof(double), hipMemcpyDeviceToDevice, st), "copy b");
  eng_.nrm2_async(b_.p, nrows, norms_.p + 0);
  eng_.solve_async<double>(x, 1, L);
  double best = INFINITY;
  int stagnant = 0;
  for (int it = 0;; ++it) {
    eng_.residual_async(x, b_.p, r_.p);
    eng_.nrm2_async(r_.p, nrows, norms_.p + 1);
    double h[2];
    hip_check(hipMemcpyAsync(h, norms_.p, 2 * sizeof(double), hipMemcpyDeviceToHost, st), "norms");
    hip_check(hipStreamSynchronize(st), "refine sync");
    const double nb = h[0] > 0 ? h[0] : 1.0;
    const double rel = h[1] / nb;
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
    // the residual becomes the correction right-hand side: r_ has nrows
    // entries; rows beyond (when ncols > nrows) are zero
    if (L > nrows) hip_check(launch_convert<double, double>(nullptr, r_.p + nrows, L - nrows, st), "pad");
    eng_.solve_async<double>(r_.p, 1, L);
    hip_check(launch_axpy(1.0, r_.p, x, d.ncols(), st), "x += d");
  }
  return rep;
}

}  // namespace bcr
