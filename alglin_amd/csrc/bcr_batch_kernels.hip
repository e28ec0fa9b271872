// Kernels used only by the batched engine: per-system preparation, partial
// sums, norms, and the synthetic block generator.
//
//   prep_batched      blocks the first reduction level does not touch (row
//                     N-1 when N is odd, C_N) copied into the working factors
//   reduce_batched    dst_s[i] += sum_{k=first}^{first+count-1} part_s[k*m+i],
//                     fixed summation order (deterministic)
//   norms_batched     out[2s] = ||b_s||_2, out[2s+1] = ||r_s||_2
//   gen_batched       counter-based Gaussian blocks (splitmix64 hash of
//                     (seed, system, kind, index), Box-Muller)
#include "bcr_batch_kernels.hh"

namespace bcr {
namespace {

template <typename TA, typename TW>
__global__ void __launch_bounds__(256) prep_batched(const PrepArgs<TA, TW>* args) {
  const PrepArgs<TA, TW>& a = args[blockIdx.x];
  const int N = a.N, n = a.n, nr = a.nr, nx = a.nx;
  const size_t nn = size_t(n) * n, rn = size_t(nr) * n, bn = size_t(n) * nx;
  if (N % 2 == 1) {
    const size_t k = size_t(N - 1);
    for (size_t i = threadIdx.x; i < nn; i += blockDim.x) {
      a.wD[k * nn + i] = TW(a.D[k * nn + i]);
      a.wE[k * nn + i] = TW(a.E[k * nn + i]);
    }
    for (size_t i = threadIdx.x; i < bn; i += blockDim.x) a.wB[k * bn + i] = TW(a.B[k * bn + i]);
    for (size_t i = threadIdx.x; i < rn; i += blockDim.x) a.wC[k * rn + i] = TW(a.C[k * rn + i]);
  }
  for (size_t i = threadIdx.x; i < rn; i += blockDim.x) a.wC[size_t(N) * rn + i] = TW(a.C[size_t(N) * rn + i]);
}

template <typename TP, typename TD>
__global__ void __launch_bounds__(256) reduce_batched(const ReduceArgs<TP, TD>* args) {
  __shared__ double red[256];
  const ReduceArgs<TP, TD>& a = args[blockIdx.x];
  for (int i = 0; i < a.m; ++i) {
    double s = 0;
    for (int k = threadIdx.x; k < a.count; k += blockDim.x) s += double(a.part[size_t(a.first + k) * a.m + i]);
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if (int(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      TD* d = a.dst + (i % a.mrow) + size_t(i / a.mrow) * a.ldd;
      *d = TD(double(*d) + red[0]);
    }
    __syncthreads();
  }
}

__global__ void __launch_bounds__(256) norms_batched(const NormArgs* args, double* out) {
  __shared__ double red[2][256];
  const NormArgs& a = args[blockIdx.x];
  double sb = 0, sr = 0;
  for (int64_t i = threadIdx.x; i < a.len; i += blockDim.x) {
    const double vb = a.b[i], vr = a.r[i];
    sb += vb * vb;
    sr += vr * vr;
  }
  red[0][threadIdx.x] = sb;
  red[1][threadIdx.x] = sr;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (int(threadIdx.x) < w) {
      red[0][threadIdx.x] += red[0][threadIdx.x + w];
      red[1][threadIdx.x] += red[1][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[2 * blockIdx.x] = sqrt(red[0][0]);
    out[2 * blockIdx.x + 1] = sqrt(red[1][0]);
  }
}

__device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ double gauss(uint64_t seed, uint64_t ctr) {
  const uint64_t z1 = splitmix64(seed ^ splitmix64(ctr));
  const uint64_t z2 = splitmix64(z1);
  const double u1 = (double((z1 >> 11) + 1)) * 0x1.0p-53;   // (0, 1]
  const double u2 = double(z2 >> 11) * 0x1.0p-53;           // [0, 1)
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

// grid (chunks, nsys); kind 0 D, 1 E, 2 B, 3 C, 4 H, 5 Cq, 6 F
__global__ void __launch_bounds__(256) gen_batched(const GenArgs* args, int kind, uint64_t seed) {
  const GenArgs& a = args[blockIdx.y];
  const int n = a.n;
  int64_t len;
  double* dst;
  switch (kind) {
    case 0: len = int64_t(a.N) * n * n; dst = a.D; break;
    case 1: len = int64_t(a.N) * n * n; dst = a.E; break;
    case 2: len = int64_t(a.N) * n * a.nx; dst = a.B; break;
    case 3: len = int64_t(a.N + 1) * a.nr * n; dst = a.C; break;
    case 4: len = int64_t(n + a.qr) * (2 * n + a.qx + a.nx); dst = a.H; break;
    case 5: len = int64_t(a.nr) * a.qx; dst = a.Cq; break;
    default: len = int64_t(a.nr) * a.nx; dst = a.F; break;
  }
  const double sc = (kind <= 1) ? 1.0 / sqrt(double(n)) : 1.0;
  const uint64_t base = ((uint64_t(a.sys) << 4) | uint64_t(kind)) << 40;
  const int64_t nn = int64_t(n) * n;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < len; i += int64_t(gridDim.x) * blockDim.x) {
    double v = sc * gauss(seed, base | uint64_t(i));
    if (kind == 0) {
      const int64_t w = i % nn;
      if (w % n == w / n) v += a.shift;
    }
    dst[i] = v;
  }
}

}  // namespace

template <typename TA, typename TW>
hipError_t launch_prep_batched(const PrepArgs<TA, TW>* args, int nsys, hipStream_t st) {
  hipLaunchKernelGGL((prep_batched<TA, TW>), dim3(nsys), dim3(256), 0, st, args);
  return hipGetLastError();
}

template <typename TP, typename TD>
hipError_t launch_reduce_batched(const ReduceArgs<TP, TD>* args, int nsys, hipStream_t st) {
  hipLaunchKernelGGL((reduce_batched<TP, TD>), dim3(nsys), dim3(256), 0, st, args);
  return hipGetLastError();
}

hipError_t launch_norms_batched(const NormArgs* args, int nsys, double* out, hipStream_t st) {
  hipLaunchKernelGGL(norms_batched, dim3(nsys), dim3(256), 0, st, args, out);
  return hipGetLastError();
}

hipError_t launch_gen_batched(const GenArgs* args, int nsys, uint64_t seed, hipStream_t st) {
  for (int kind = 0; kind < 7; ++kind) {
    hipLaunchKernelGGL(gen_batched, dim3(kind <= 1 ? 64 : 8, nsys), dim3(256), 0, st, args, kind, seed);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

template hipError_t launch_prep_batched<double, double>(const PrepArgs<double, double>*, int, hipStream_t);
template hipError_t launch_prep_batched<double, float>(const PrepArgs<double, float>*, int, hipStream_t);
template hipError_t launch_reduce_batched<double, double>(const ReduceArgs<double, double>*, int, hipStream_t);
template hipError_t launch_reduce_batched<float, double>(const ReduceArgs<float, double>*, int, hipStream_t);

}  // namespace bcr
