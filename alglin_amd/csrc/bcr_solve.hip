// Solve-phase kernels of the bordered cyclic reduction.
//
//   forward_kernel   one level of the forward sweep (cc:1615-1719): replays the
//                    pair transform on (x_jp, x_j) and writes the border partial
//                    -C_j' x_j to Xpart[j]
//   backward_kernel  one level of the back substitution (cc:1805-1949):
//                    x_j = U^{-1}(x_j - D_j x_jp - E_j x_jpp - B_j xb), QRP swaps
//   reduce_add       dst += sum_j part[j] (deterministic two-pass tree)
//
// One workgroup per pair; the pair's factor blocks are staged into LDS with
// coalesced loads (all in flight at once), the short sequential recurrences
// then run out of LDS.  The right-hand-side block of the pair is processed in
// chunks of up to kRhsChunk columns.
#include "bcr_launch.hh"

namespace bcr {


template <typename TW, typename TX, int METHOD>
__device__ __forceinline__ void forward_body(SolveArgs<TW, TX> a, int bid) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int p = bid;
  if (p >= a.npairs) return;
  int j, jp, jpp;
  pair_of_uniform(a, p, j, jp, jpp);
  const int n = a.n, n2 = 2 * n, nr = a.nr;
  const int tid = threadIdx.x, nth = blockDim.x;
  TW* Tp = reinterpret_cast<TW*>(smem_raw);                                     // 2n*n + n
  TX* X = reinterpret_cast<TX*>(smem_raw + align16(tsize(n) * sizeof(TW)));   // 2n*kRhsChunk
  __shared__ int s_perm[64];

  const TW* Tj = a.Tm + size_t(j) * tsize(n);
  const int tcount = (METHOD == kLU) ? n2 * n : int(tsize(n));
  for (int i = tid; i < tcount; i += nth) Tp[i] = Tj[i];
  for (int i = tid; i < n; i += nth) s_perm[i] = a.perm[size_t(j) * n + i];

  for (int q0 = 0; q0 < a.nrhs; q0 += kRhsChunk) {
    const int nq = min(kRhsChunk, a.nrhs - q0);
    for (int i = tid; i < n2 * nq; i += nth) {
      const int r = i % n2, q = i / n2;
      const size_t row = r < n ? size_t(jp) * n + r : size_t(j) * n + (r - n);
      X[r + q * n2] = a.x[row + size_t(q0 + q) * a.ldx];
    }
    __syncthreads();
    if (METHOD == kLU) {
      // row interchanges first (laswp on rows 1..n, cc:939), then L^{-1}
      for (int q = tid; q < nq; q += nth)
        for (int c = 0; c < n; ++c) {
          const int pv = s_perm[c] - 1;
          if (pv != c) { TX t = X[c + q * n2]; X[c + q * n2] = X[pv + q * n2]; X[pv + q * n2] = t; }
        }
      __syncthreads();
      for (int c = 0; c < n; ++c) {
        for (int i = tid; i < (n2 - c - 1) * nq; i += nth) {
          const int r = c + 1 + i % (n2 - c - 1), q = i / (n2 - c - 1);
          X[r + q * n2] -= TX(Tp[r + c * n2]) * X[c + q * n2];
        }
        __syncthreads();
      }
    } else {
      // Q^T = H_{n-1} ... H_0 (ormqr 'L','T', cc:962-971): one wave per
      // right-hand side, its rows lane and lane + 64 (2n <= 128) in
      // registers; per reflector a wave dot product (a butterfly: every
      // lane ends with the same sum) and an axpy, no barrier
      const int lane = tid & 63, wv = tid >> 6, nw = nth >> 6;
      for (int q = wv; q < nq; q += nw) {
        TX* xq = X + q * n2;
        TX x0 = lane < n2 ? xq[lane] : TX(0);
        TX x1 = lane + 64 < n2 ? xq[lane + 64] : TX(0);
        for (int c = 0; c < n; ++c) {
          const TX tau = TX(Tp[size_t(n2) * n + c]);
          // v = (1 at row c, the stored reflector below it); row c < 64
          const TX v0 = lane == c ? TX(1) : (lane > c && lane < n2 ? TX(Tp[lane + c * n2]) : TX(0));
          const TX v1 = lane + 64 < n2 ? TX(Tp[lane + 64 + c * n2]) : TX(0);
          TX d = fma(v0, x0, v1 * x1);
#pragma unroll
          for (int off = 32; off >= 1; off >>= 1) d += __shfl_xor(d, off);
          const TX w = d * tau;
          x0 = fma(-v0, w, x0);
          x1 = fma(-v1, w, x1);
        }
        if (lane < n2) xq[lane] = x0;
        if (lane + 64 < n2) xq[lane + 64] = x1;
      }
      __syncthreads();
    }
    // TOP (x_jp) <- W[n:2n], BOTTOM (x_j) <- W[0:n]   (cc:1041-1042)
    for (int i = tid; i < n2 * nq; i += nth) {
      const int r = i % n2, q = i / n2;
      const size_t row = r < n ? size_t(j) * n + r : size_t(jp) * n + (r - n);
      a.x[row + size_t(q0 + q) * a.ldx] = X[r + q * n2];
    }
    // border partial: Xpart[j] = -C_j' x_j   (cc:1644-1650)
    const TW* Cj = a.C + size_t(j) * nr * n;
    for (int i = tid; i < nr * nq; i += nth) {
      const int r = i % nr, q = i / nr;
      TX s = 0;
      for (int c = 0; c < n; ++c) s += TX(Cj[r + c * nr]) * X[c + q * n2];
      a.Xpart[size_t(j) * nr * a.nrhs + r + size_t(q0 + q) * nr] = -s;
    }
    __syncthreads();
  }
}

template <typename TW, typename TX, int METHOD>
__global__ void __launch_bounds__(128) forward_kernel(SolveArgs<TW, TX> a) {
  forward_body<TW, TX, METHOD>(a, blockIdx.x);
}


template <typename TW, typename TX, int METHOD>
__device__ __forceinline__ void backward_body(SolveArgs<TW, TX> a, int bid) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int p = bid;
  if (p >= a.npairs) return;
  int j, jp, jpp;
  pair_of_uniform(a, p, j, jp, jpp);
  const int n = a.n, n2 = 2 * n, nx = a.nx;
  const int tid = threadIdx.x, nth = blockDim.x;
  const size_t nn = size_t(n) * n;
  TW* sD = reinterpret_cast<TW*>(smem_raw);
  TW* sE = sD + nn;
  TW* sU = sE + nn;
  TW* sB = sU + nn;
  TX* Xs = reinterpret_cast<TX*>(smem_raw + align16((3 * nn + size_t(n) * nx) * sizeof(TW)));  // inputs
  TX* Y = Xs + size_t(3 * n + nx) * kRhsChunk;
  __shared__ int s_perm[64];

  const TW* Tj = a.Tm + size_t(j) * tsize(n);
  for (int i = tid; i < n * n; i += nth) {
    sD[i] = a.D[size_t(j) * nn + i];
    sE[i] = a.E[size_t(j) * nn + i];
    const int r = i % n, c = i / n;
    sU[i] = Tj[r + c * n2];
  }
  for (int i = tid; i < n * nx; i += nth) sB[i] = a.B[size_t(j) * n * nx + i];
  if (METHOD == kQRP)
    for (int i = tid; i < n; i += nth) s_perm[i] = a.perm[size_t(j) * n + i];
  const size_t xbrow = size_t(a.N + 1) * n + a.qx;   // cc:1812
  const int W = 3 * n + nx;

  for (int q0 = 0; q0 < a.nrhs; q0 += kRhsChunk) {
    const int nq = min(kRhsChunk, a.nrhs - q0);
    for (int i = tid; i < W * nq; i += nth) {
      const int r = i % W, q = i / W;
      size_t row;
      if (r < n) row = size_t(j) * n + r;
      else if (r < 2 * n) row = size_t(jp) * n + (r - n);
      else if (r < 3 * n) row = size_t(jpp) * n + (r - 2 * n);
      else row = xbrow + (r - 3 * n);
      Xs[r + q * W] = a.x[row + size_t(q0 + q) * a.ldx];
    }
    __syncthreads();
    for (int i = tid; i < n * nq; i += nth) {
      const int r = i % n, q = i / n;
      const TX* xq = Xs + q * W;
      TX acc = xq[r];
      for (int c = 0; c < n; ++c) acc -= TX(sD[r + c * n]) * xq[n + c] + TX(sE[r + c * n]) * xq[2 * n + c];
      for (int c = 0; c < nx; ++c) acc -= TX(sB[r + c * n]) * xq[3 * n + c];
      Xs[r + q * W] = acc;
    }
    __syncthreads();
    // U y = acc, column-oriented; y never overwrites the row it reads
    for (int c = n - 1; c >= 0; --c) {
      const TX uinv = TX(1) / TX(sU[c + c * n]);
      for (int i = tid; i < (c + 1) * nq; i += nth) {
        const int r = i % (c + 1), q = i / (c + 1);
        const TX yc = Xs[c + q * W] * uinv;
        if (r == c) Y[c + q * n] = yc;
        else Xs[r + q * W] -= TX(sU[r + c * n]) * yc;
      }
      __syncthreads();
    }
    if (METHOD == kQRP) {   // cc:1851-1856
      for (int q = tid; q < nq; q += nth)
        for (int i = 0; i < n; ++i) {
          const int s = s_perm[i];
          if (s > i) { TX t = Y[i + q * n]; Y[i + q * n] = Y[s + q * n]; Y[s + q * n] = t; }
        }
      __syncthreads();
    }
    for (int i = tid; i < n * nq; i += nth) {
      const int r = i % n, q = i / n;
      a.x[size_t(j) * n + r + size_t(q0 + q) * a.ldx] = Y[r + q * n];
    }
    __syncthreads();
  }
}

template <typename TW, typename TX, int METHOD>
__global__ void __launch_bounds__(128) backward_kernel(SolveArgs<TW, TX> a) {
  backward_body<TW, TX, METHOD>(a, blockIdx.x);
}


// part: [count][m] (column j of the partial sums), dst[i] += sum_j part[j][i],
// dst index i maps to dst[(i % mrow) + (i / mrow) * ldd].
template <typename TP, typename TD>
__global__ void __launch_bounds__(256) reduce_partial_kernel(const TP* __restrict__ part, int first, int count,
                                                             int m, TP* __restrict__ tmp, int nchunks) {
  // pass 1: chunk b sums rows [lo, hi) of part for every i < m
  const int b = blockIdx.x;
  const int per = (count + nchunks - 1) / nchunks;
  const int lo = first + b * per, hi = min(first + count, lo + per);
  for (int i = threadIdx.x; i < m; i += blockDim.x) {
    TP s = 0;
    for (int k = lo; k < hi; ++k) s += part[size_t(k) * m + i];
    tmp[size_t(b) * m + i] = s;
  }
}

// one wave per output element: lanes stride over the chunk partials, then a
// wave sum (deterministic order)
template <typename TP, typename TD>
__global__ void __launch_bounds__(64) reduce_final_kernel(const TP* __restrict__ tmp, int nchunks, int m,
                                                          TD* __restrict__ dst, int mrow, int64_t ldd) {
  const int i = blockIdx.x;
  if (i >= m) return;
  TP s = 0;
  for (int b = threadIdx.x; b < nchunks; b += 64) s += tmp[size_t(b) * m + i];
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off);
  if (threadIdx.x == 0) dst[(i % mrow) + size_t(i / mrow) * ldd] += TD(s);
}

template <typename TW, typename TX>
hipError_t launch_forward(int method, SolveArgs<TW, TX> const& a, hipStream_t st) {
  const size_t bytes = align16(tsize(a.n) * sizeof(TW)) + size_t(2 * a.n) * kRhsChunk * sizeof(TX);
  dim3 grid(a.npairs), block(a.n <= 32 ? 64 : 128);
  if (method == kLU) hipLaunchKernelGGL((forward_kernel<TW, TX, kLU>), grid, block, bytes, st, a);
  else hipLaunchKernelGGL((forward_kernel<TW, TX, kQR>), grid, block, bytes, st, a);
  return hipGetLastError();
}

template <typename TW, typename TX>
hipError_t launch_backward(int method, SolveArgs<TW, TX> const& a, hipStream_t st) {
  const size_t nn = size_t(a.n) * a.n;
  const size_t bytes = align16((3 * nn + size_t(a.n) * a.nx) * sizeof(TW)) +
                       (size_t(3 * a.n + a.nx) + a.n) * kRhsChunk * sizeof(TX);
  dim3 grid(a.npairs), block(a.n <= 32 ? 64 : 128);
  if (method == kQRP) hipLaunchKernelGGL((backward_kernel<TW, TX, kQRP>), grid, block, bytes, st, a);
  else hipLaunchKernelGGL((backward_kernel<TW, TX, kLU>), grid, block, bytes, st, a);
  return hipGetLastError();
}

template <typename TW, typename TX, int METHOD, bool FWD>
__global__ void __launch_bounds__(128) sweep_generic_batched(BatchTable<SolveArgs<TW, TX>> t) {
  const int e = t.entry[blockIdx.x];
  if (skipped(t.args[e].skip)) return;
  if (FWD) forward_body<TW, TX, METHOD>(t.args[e], int(blockIdx.x) - t.first[e]);
  else backward_body<TW, TX, METHOD>(t.args[e], int(blockIdx.x) - t.first[e]);
}

// one workgroup per pair; n is the largest block size of the launch
template <typename TW, typename TX>
hipError_t launch_sweep_generic_batched(bool forward, int method, BatchTable<SolveArgs<TW, TX>> const& t, int nblocks,
                                        int n, int nx, hipStream_t st) {
  const size_t nn = size_t(n) * n;
  const size_t fbytes = align16(tsize(n) * sizeof(TW)) + size_t(2 * n) * kRhsChunk * sizeof(TX);
  const size_t bbytes = align16((3 * nn + size_t(n) * nx) * sizeof(TW)) + (size_t(3 * n + nx) + n) * kRhsChunk * sizeof(TX);
  dim3 grid(nblocks), block(n <= 32 ? 64 : 128);
  if (forward) {
    if (method == kLU) hipLaunchKernelGGL((sweep_generic_batched<TW, TX, kLU, true>), grid, block, fbytes, st, t);
    else hipLaunchKernelGGL((sweep_generic_batched<TW, TX, kQR, true>), grid, block, fbytes, st, t);
  } else {
    if (method == kQRP) hipLaunchKernelGGL((sweep_generic_batched<TW, TX, kQRP, false>), grid, block, bbytes, st, t);
    else hipLaunchKernelGGL((sweep_generic_batched<TW, TX, kLU, false>), grid, block, bbytes, st, t);
  }
  return hipGetLastError();
}

template hipError_t launch_sweep_generic_batched<double, double>(bool, int, BatchTable<SolveArgs<double, double>> const&,
                                                                 int, int, int, hipStream_t);
template hipError_t launch_sweep_generic_batched<float, float>(bool, int, BatchTable<SolveArgs<float, float>> const&,
                                                               int, int, int, hipStream_t);
template hipError_t launch_sweep_generic_batched<float, double>(bool, int, BatchTable<SolveArgs<float, double>> const&,
                                                                int, int, int, hipStream_t);

// dst[(i%mrow) + (i/mrow)*ldd] += sum_{k=first}^{first+count-1} part[k*m + i]
template <typename TP, typename TD>
hipError_t launch_reduce(const TP* part, int first, int count, int m, TP* tmp, int max_chunks, TD* dst, int mrow,
                         int64_t ldd, hipStream_t st) {
  if (m <= 0 || count <= 0) return hipSuccess;
  int nchunks = min(max_chunks, max(1, count / 16));
  hipLaunchKernelGGL((reduce_partial_kernel<TP, TD>), dim3(nchunks), dim3(256), 0, st, part, first, count, m, tmp,
                     nchunks);
  hipLaunchKernelGGL((reduce_final_kernel<TP, TD>), dim3(m), dim3(64), 0, st, tmp, nchunks, m, dst, mrow, ldd);
  return hipGetLastError();
}

template hipError_t launch_forward<double, double>(int, SolveArgs<double, double> const&, hipStream_t);
template hipError_t launch_forward<float, float>(int, SolveArgs<float, float> const&, hipStream_t);
template hipError_t launch_forward<float, double>(int, SolveArgs<float, double> const&, hipStream_t);
template hipError_t launch_backward<double, double>(int, SolveArgs<double, double> const&, hipStream_t);
template hipError_t launch_backward<float, float>(int, SolveArgs<float, float> const&, hipStream_t);
template hipError_t launch_backward<float, double>(int, SolveArgs<float, double> const&, hipStream_t);
template hipError_t launch_reduce<double, double>(const double*, int, int, int, double*, int, double*, int, int64_t,
                                                  hipStream_t);
template hipError_t launch_reduce<float, float>(const float*, int, int, int, float*, int, float*, int, int64_t,
                                                hipStream_t);
template hipError_t launch_reduce<float, double>(const float*, int, int, int, float*, int, double*, int, int64_t,
                                                 hipStream_t);
template hipError_t launch_reduce<double, float>(const double*, int, int, int, double*, int, float*, int, int64_t,
                                                 hipStream_t);

}  // namespace bcr
