// Singular-value last block (select_last SVD / LSS / LSY / PINV, and the
// second solver of BABD_BorderedCR.cc:1347-1369 when the first fails), FP64,
// one workgroup.
//
// One-sided Jacobi on W (rows x cols, rows = max(Nr, Nc), W = A or A^T):
// column pairs are rotated until every pair is orthogonal to eps relative,
// visited in round-robin order (cols/2 disjoint pairs per round, one
// wavefront per pair, cols-1 rounds per sweep), at most 80 sweeps.  Then the
// singular values s_j = ||W_j||, sorted descending; values at or below
// max(Nr, Nc) * eps * s_max count as zero.  The solution of every variant is
// the minimum-norm least-squares one, x = V S^+ U^T b (the dense SVD / LSS /
// LSY / PINV classes all return it for the square or tall blocks here).
// Outputs: U (Nr x cols), V (Nc x cols), S (cols), rank; rank 0 sets status
// bit 2 (failed factorization).
//
// As the second solver: svd runs only when the first set status bit 2; it
// clears that bit and sets bit 3, which the solve reads to pick this path.
#include "bcr_launch.hh"

namespace bcr {
namespace {

constexpr int kSvdMax = 512;   // cols bound of the static tables

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

__global__ void __launch_bounds__(1024) svd_last_kernel(SvdArgs a) {
  __shared__ double s_val[kSvdMax];
  __shared__ int s_ord[kSvdMax];
  __shared__ double s_off[16];
  __shared__ int s_go;
  const int tid = threadIdx.x, nth = blockDim.x, lane = tid & 63, wid = tid >> 6, nw = nth >> 6;
  if (a.conditional) {
    if (tid == 0) s_go = (*a.status & 2) ? 1 : 0;
    __syncthreads();
    if (!s_go) return;
    __syncthreads();
    if (tid == 0) *a.status = (*a.status & ~2) | 8;
  }
  const int m = a.Nr, n = a.Nc;
  const bool tr = m < n;
  const int rows = tr ? n : m, cols = tr ? m : n;
  double* W = a.W;    // rows x cols
  double* Vj = a.Vj;  // cols x cols
  for (int i = tid; i < rows * cols; i += nth) {
    const int r = i % rows, c = i / rows;
    W[i] = tr ? a.A[c + size_t(r) * m] : a.A[r + size_t(c) * m];
  }
  for (int i = tid; i < cols * cols; i += nth) Vj[i] = (i % cols == i / cols) ? 1.0 : 0.0;
  __syncthreads();
  const double eps = 2.220446049250313e-16;
  const int P = cols + (cols & 1);   // players of the round robin (one dummy when odd)
  for (int sweep = 0; sweep < 80; ++sweep) {
    double off = 0;
    for (int round = 0; round < P - 1; ++round) {
      for (int k = wid; k < P / 2; k += nw) {
        // circle method: player 0 fixed, players 1..P-1 rotate by `round`
        int p = k == 0 ? 0 : 1 + (k - 1 + round) % (P - 1);
        int q = 1 + (P - 2 - k + round) % (P - 1);
        if (p > q) { const int t = p; p = q; q = t; }
        if (q >= cols) continue;   // the dummy player
        double al = 0, be = 0, ga = 0;
        for (int i = lane; i < rows; i += 64) {
          const double wp = W[i + size_t(p) * rows], wq = W[i + size_t(q) * rows];
          al += wp * wp;
          be += wq * wq;
          ga += wp * wq;
        }
        al = wsum(al);
        be = wsum(be);
        ga = wsum(ga);
        if (ga == 0.0) continue;
        const double rel = fabs(ga) / sqrt(al * be);
        off = fmax(off, rel);
        if (rel <= eps) continue;
        const double zeta = (be - al) / (2.0 * ga);
        const double t = copysign(1.0, zeta) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
        const double c = 1.0 / sqrt(1.0 + t * t), s = c * t;
        for (int i = lane; i < rows; i += 64) {
          const double wp = W[i + size_t(p) * rows], wq = W[i + size_t(q) * rows];
          W[i + size_t(p) * rows] = c * wp - s * wq;
          W[i + size_t(q) * rows] = s * wp + c * wq;
        }
        for (int i = lane; i < cols; i += 64) {
          const double vp = Vj[i + size_t(p) * cols], vq = Vj[i + size_t(q) * cols];
          Vj[i + size_t(p) * cols] = c * vp - s * vq;
          Vj[i + size_t(q) * cols] = s * vp + c * vq;
        }
      }
      __threadfence_block();
      __syncthreads();
    }
    if (lane == 0) s_off[wid] = off;
    __syncthreads();
    double o = 0;
    for (int w = 0; w < nw; ++w) o = fmax(o, s_off[w]);
    __syncthreads();
    if (o <= eps) break;
  }
  // singular values, sorted descending (stable)
  for (int j = wid; j < cols; j += nw) {
    double ss = 0;
    for (int i = lane; i < rows; i += 64) ss += W[i + size_t(j) * rows] * W[i + size_t(j) * rows];
    ss = wsum(ss);
    if (lane == 0) s_val[j] = sqrt(ss);
  }
  __syncthreads();
  if (tid == 0) {
    for (int j = 0; j < cols; ++j) s_ord[j] = j;
    for (int j = 1; j < cols; ++j) {   // insertion sort, stable
      const int v = s_ord[j];
      int i = j - 1;
      while (i >= 0 && s_val[s_ord[i]] < s_val[v]) { s_ord[i + 1] = s_ord[i]; --i; }
      s_ord[i + 1] = v;
    }
    const double smax = cols > 0 ? s_val[s_ord[0]] : 0.0;
    const double tol = double(m > n ? m : n) * eps * smax;
    int rank = 0;
    for (int l = 0; l < cols; ++l) {
      const double sv = s_val[s_ord[l]];
      a.S[l] = sv;
      if (sv > tol && sv > 0.0) ++rank;
    }
    *a.rank = rank;
    if (rank == 0) atomicOr(a.status, 2);
  }
  __syncthreads();
  // U (m x cols) and V (n x cols) in sorted order
  for (int i = tid; i < rows * cols; i += nth) {
    const int r = i % rows, l = i / rows, j = s_ord[l];
    const double sv = s_val[j];
    const double u = sv > 0.0 ? W[r + size_t(j) * rows] / sv : 0.0;
    if (tr) a.V[r + size_t(l) * n] = u;
    else a.U[r + size_t(l) * m] = u;
  }
  for (int i = tid; i < cols * cols; i += nth) {
    const int r = i % cols, l = i / cols, j = s_ord[l];
    const double v = Vj[r + size_t(j) * cols];
    if (tr) a.U[r + size_t(l) * m] = v;
    else a.V[r + size_t(l) * n] = v;
  }
}

}  // namespace

hipError_t launch_svd_last(SvdArgs const& a, hipStream_t st) {
  const int cols = a.Nr < a.Nc ? a.Nr : a.Nc;
  if (cols > kSvdMax) return hipErrorInvalidValue;
  hipLaunchKernelGGL(svd_last_kernel, dim3(1), dim3(1024), 0, st, a);
  return hipGetLastError();
}

}  // namespace bcr
