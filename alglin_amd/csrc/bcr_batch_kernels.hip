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
//   panel_stats_batched, features_batched
//                     the selector's features: per-panel pivot statistics
//                     (one wave per panel), then one workgroup per system
#include "bcr_batch_kernels.hh"
#include "bcr_wave.hh"

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
  if (skipped(a.skip)) return;
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
  if (skipped(a.skip)) return;
  double sb = 0, sr = 0;
  // four elements' loads in flight per thread; each thread still sums its
  // elements i, i + blockDim, i + 2 blockDim, ... in that order
  const int64_t st = blockDim.x;
  int64_t i = threadIdx.x;
  for (; i + 3 * st < a.len; i += 4 * st) {
    double vb[4], vr[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      vb[u] = a.b[i + u * st];
      vr[u] = a.r[i + u * st];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      sb += vb[u] * vb[u];
      sr += vr[u] * vr[u];
    }
  }
  for (; i < a.len; i += st) {
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
    out[2 * a.sys] = sqrt(red[0][0]);
    out[2 * a.sys + 1] = sqrt(red[1][0]);
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
  // the block of this kind that belongs to the rank-deficient pair column
  const int64_t sblk = a.sing_k < 1 ? -1 : kind == 0 ? a.sing_k : kind == 1 ? a.sing_k - 1 : -1;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < len; i += int64_t(gridDim.x) * blockDim.x) {
    double v = sc * gauss(seed, base | uint64_t(i));
    int64_t w = 0;
    if (kind <= 1) w = i % nn;
    if (kind == 0 && w % n == w / n) v += a.shift;
    if (kind <= 1 && n >= 3 && i / nn == sblk && w / n == n - 1) {
      const int64_t r = w % n, b0 = (i / nn) * nn;
      double c0 = sc * gauss(seed, base | uint64_t(b0 + r));
      double c1 = sc * gauss(seed, base | uint64_t(b0 + n + r));
      if (kind == 0) {
        if (r == 0) c0 += a.shift;
        if (r == 1) c1 += a.shift;
      }
      v = 0.5 * (c0 + c1) + a.eps * sc * gauss(seed, base | uint64_t(i));
    }
    dst[i] = v;
  }
}

// one wave per pair panel (four per workgroup, every system in one launch):
// the lanes walk the top n x n block of T (U on and above the diagonal) in
// flat order, four loads in flight per lane, and write the panel's min and
// max pivot magnitude and its largest |u_ik| (i <= k); the wave reductions
// run on the float bit patterns (non-negative floats order as unsigned)
template <typename TW>
__global__ void __launch_bounds__(256) panel_stats_batched(BatchTable<FeatArgs<TW>> t) {
  const int e = t.entry[blockIdx.x];
  const FeatArgs<TW>& a = t.args[e];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int j = 1 + 4 * (int(blockIdx.x) - t.first[e]) + w;
  if (j >= a.N) return;   // no barrier in this kernel
  const int n = a.n, n2 = 2 * n, nn = n * n;
  const TW* T = a.Tm + size_t(j) * tsize(n);
  const int dr = 64 % n, dc = 64 / n;
  int row = lane % n, col = lane / n;
  float dmin = INFINITY, dmax = 0.f, umax = 0.f;
  for (int k0 = 0; k0 < nn; k0 += 256) {
    float v[4];
    int rr[4], cc[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      rr[u] = row;
      cc[u] = col;
      v[u] = keep_if(float(fabs(T[row + size_t(col < n ? col : n - 1) * n2])), col < n);
      row += dr;
      col += dc;
      if (row >= n) {
        row -= n;
        ++col;
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (cc[u] < n && rr[u] == cc[u]) {
        dmin = fminf(dmin, v[u]);
        dmax = fmaxf(dmax, v[u]);
      }
      if (cc[u] < n && rr[u] <= cc[u]) umax = fmaxf(umax, v[u]);
    }
  }
  const unsigned pmin = wave::min_u32(__float_as_uint(dmin));
  const unsigned pmax = wave::max_u32(__float_as_uint(dmax));
  const unsigned um = wave::max_u32(__float_as_uint(umax));
  if (lane == 0) {
    float* o = a.pstat + 3 * size_t(j - 1);
    o[0] = __uint_as_float(pmin);
    o[1] = __uint_as_float(pmax);
    o[2] = __uint_as_float(um);
  }
}

// one workgroup per system: the panels' triples (coalesced), a block
// reduction, and the 12 features
template <typename TW>
__global__ void __launch_bounds__(256) features_batched(const FeatArgs<TW>* args, float* out) {
  __shared__ double r_min[256], r_max[256], r_sum[256], r_rat[256], r_cnt[256], r_up[256];
  const FeatArgs<TW>& a = args[blockIdx.x];
  const int t = threadIdx.x;
  double dmin = INFINITY, dmax = 0, lsum = 0, rmin = INFINITY, cnt = 0, umax = 0;
  for (int i = t; i < a.N - 1; i += 256) {
    const float* ps = a.pstat + 3 * size_t(i);
    const double pmin = ps[0], pmax = ps[1], pu = ps[2];
    const double ratio = pmax > 0 ? pmin / pmax : 0.0;
    dmin = fmin(dmin, pmin);
    dmax = fmax(dmax, pmax);
    lsum += log10(ratio + 1e-30);
    rmin = fmin(rmin, ratio);
    cnt += ratio < 1e-5 ? 1.0 : 0.0;
    umax = fmax(umax, pu);
  }
  r_min[t] = dmin; r_max[t] = dmax; r_sum[t] = lsum; r_rat[t] = rmin; r_cnt[t] = cnt; r_up[t] = umax;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (t < h) {
      r_min[t] = fmin(r_min[t], r_min[t + h]);
      r_max[t] = fmax(r_max[t], r_max[t + h]);
      r_sum[t] += r_sum[t + h];
      r_rat[t] = fmin(r_rat[t], r_rat[t + h]);
      r_cnt[t] += r_cnt[t + h];
      r_up[t] = fmax(r_up[t], r_up[t + h]);
    }
    __syncthreads();
  }
  if (t == 0) {
    const int n = a.n;
    float* f = out + size_t(a.sys) * 12;
    const int np = a.N - 1;
    f[0] = float(log2(double(a.N)) / 12.0);
    f[1] = float(n) / 48.f;
    f[2] = float(a.qr) / 12.f;
    f[3] = float(a.nr) / 8.f;
    if (np > 0) {
      f[4] = float(log10(r_min[0] + 1e-30));
      f[5] = float(log10(r_max[0] + 1e-30));
      f[6] = float(r_sum[0] / np);
      f[7] = float(log10(r_rat[0] + 1e-30));
      f[8] = float(r_cnt[0] / np);
      f[11] = float(log10(r_up[0] / (r_max[0] + 1e-30) + 1e-30));
    } else {
      f[4] = f[5] = f[6] = f[7] = f[8] = f[11] = 0.f;
    }
    double lmin = INFINITY, lmax = 0;
    for (int i = 0; i < a.Nc; ++i) {
      const double d = fabs(a.Hm[i + size_t(i) * a.Nc]);
      lmin = fmin(lmin, d);
      lmax = fmax(lmax, d);
    }
    f[9] = float(log10((lmax > 0 ? lmin / lmax : 0.0) + 1e-30));
    const double nb = a.norms[2 * a.sys] > 0 ? a.norms[2 * a.sys] : 1.0;
    f[10] = float(log10(a.norms[2 * a.sys + 1] / nb + 1e-30));
  }
}

__global__ void __launch_bounds__(256) copy_batched(const CopyArgs* args) {
  const CopyArgs& a = args[blockIdx.x];
  for (int64_t i = threadIdx.x; i < a.len; i += blockDim.x) a.dst[i] = a.src[i];
}

__global__ void __launch_bounds__(256) converge_batched(ConvergeArgs a) {
  const int s = int(blockIdx.x * blockDim.x + threadIdx.x);
  if (s >= a.nsys || a.done[s]) return;
  const double nb = a.norms[2 * s] > 0 ? a.norms[2 * s] : 1.0;
  const double rel = a.norms[2 * s + 1] / nb;
  const double prev = a.rel[s];
  a.rel[s] = rel;
  a.sweeps[s] = a.it;
  if (a.status[s] != 0 || !(rel == rel) || rel <= a.tol) {
    a.done[s] = 1;
    return;
  }
  const double best = a.it == 0 ? INFINITY : a.best[s];
  int stag = a.it == 0 ? 0 : a.stagnant[s];
  if (rel > 0.5 * best) {
    if (++stag >= 3) {
      a.stagnant[s] = stag;
      a.done[s] = 1;
      return;
    }
  } else {
    stag = 0;
  }
  a.best[s] = fmin(best, rel);
  a.stagnant[s] = stag;
  if (a.giveup > 0 && a.it >= 1 && prev > 0 && rel < prev) {
    // sweeps still needed at the last contraction rate
    const double need = log(a.tol / rel) / log(rel / prev);
    if (need > double(a.giveup)) {
      a.done[s] = 1;
      return;
    }
  }
  atomicAdd(a.open, 1);
}

__global__ void __launch_bounds__(256) add_batched(const CopyArgs* args) {
  const CopyArgs& a = args[blockIdx.x];
  if (skipped(a.skip)) return;
  const int64_t st = blockDim.x;
  int64_t i = threadIdx.x;
  for (; i + 3 * st < a.len; i += 4 * st) {   // four elements' loads in flight per thread
    double d[4], v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      d[u] = a.dst[i + u * st];
      v[u] = a.src[i + u * st];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) a.dst[i + u * st] = d[u] + v[u];
  }
  for (; i < a.len; i += st) a.dst[i] += a.src[i];
}

}  // namespace

template <typename TW>
hipError_t launch_panel_stats_batched(BatchTable<FeatArgs<TW>> const& t, int nblocks, hipStream_t st) {
  if (nblocks < 1) return hipSuccess;
  hipLaunchKernelGGL((panel_stats_batched<TW>), dim3(nblocks), dim3(256), 0, st, t);
  return hipGetLastError();
}

template <typename TW>
hipError_t launch_features_batched(const FeatArgs<TW>* args, int nsys, float* out, hipStream_t st) {
  hipLaunchKernelGGL((features_batched<TW>), dim3(nsys), dim3(256), 0, st, args, out);
  return hipGetLastError();
}

hipError_t launch_copy_batched(const CopyArgs* args, int nsys, hipStream_t st) {
  hipLaunchKernelGGL(copy_batched, dim3(nsys), dim3(256), 0, st, args);
  return hipGetLastError();
}

hipError_t launch_add_batched(const CopyArgs* args, int nsys, hipStream_t st) {
  hipLaunchKernelGGL(add_batched, dim3(nsys), dim3(256), 0, st, args);
  return hipGetLastError();
}

hipError_t launch_converge_batched(ConvergeArgs const& a, hipStream_t st) {
  if (a.nsys < 1) return hipSuccess;
  hipLaunchKernelGGL(converge_batched, dim3((a.nsys + 255) / 256), dim3(256), 0, st, a);
  return hipGetLastError();
}

template hipError_t launch_panel_stats_batched<float>(BatchTable<FeatArgs<float>> const&, int, hipStream_t);
template hipError_t launch_panel_stats_batched<double>(BatchTable<FeatArgs<double>> const&, int, hipStream_t);
template hipError_t launch_features_batched<float>(const FeatArgs<float>*, int, float*, hipStream_t);
template hipError_t launch_features_batched<double>(const FeatArgs<double>*, int, float*, hipStream_t);

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

namespace {
__global__ void __launch_bounds__(256) berr_batched(const NormArgs* args, double* out) {
  __shared__ double red[256];
  const NormArgs& a = args[blockIdx.x];
  if (skipped(a.skip)) return;
  double m = 0;
  for (int64_t i = threadIdx.x; i < a.len; i += blockDim.x) {
    const double ar = fabs(a.r[i]), d = a.b[i];
    const double q = ar == 0.0 ? 0.0 : (d > 0.0 ? ar / d : INFINITY);
    m = q > m || q != q ? q : m;
  }
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (int(threadIdx.x) < w) {
      const double o = red[threadIdx.x + w];
      if (o > red[threadIdx.x] || o != o) red[threadIdx.x] = o;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[a.sys] = red[0];
}
}  // namespace

hipError_t launch_berr_batched(const NormArgs* args, int nsys, double* out, hipStream_t st) {
  hipLaunchKernelGGL(berr_batched, dim3(nsys), dim3(256), 0, st, args, out);
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
