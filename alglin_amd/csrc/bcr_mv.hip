// FP64 block SpMV of the bordered BABD operator (BABD_BorderedCR.cc:2111-2166):
// r = b - A x (residual mode) or y = A x, on the pristine (unfactorized) blocks.
//
// mv_rows_kernel   one wave per block row k: the n rows D_k x_k + E_k x_{k+1}
//                  + B_k xb, and the border partial C_k x_k -> Cpart[k]
//                  (k = 0..N; row N contributes only its border partial)
// mv_tail_kernel   the n+qr boundary rows and the nr border rows
//                  (sum of the partials + Cq q + F xb)
// The per-block partials make the border rows a deterministic reduction
// instead of N serial gemv accumulations.
//
// ABS = true computes |A| |x| (+ |b| when b is given) with the same loops:
// the denominator of the componentwise backward error
// max_i |b - A x|_i / (|A| |x| + |b|)_i (Oettli-Prager) that the
// refinement uses to recognise the FP64 residual floor.
#include "bcr_launch.hh"
#include "bcr_wave.hh"

namespace bcr {


// Loads are issued eight columns at a time from clamped addresses and masked
// after (keep_if), so each chunk costs one memory latency instead of one per
// column; the sums run in the original column order.  A block row uses
// g = max(n, nr) lanes; for g <= 32 a wave takes 64 / g block rows (lane =
// sub-row * g + row), so small blocks do not leave most lanes idle; x is held one entry per lane and broadcast per
// column (v_readlane with one block row per wave, ds_bpermute within a
// sub-row group otherwise).
template <bool ABS>
__device__ __forceinline__ double mag(double v) { return ABS ? fabs(v) : v; }

template <typename TA, bool PACKED, bool ABS>
__device__ __forceinline__ void mv_rows_run(MvArgs<TA> const& a, int bid) {
  constexpr int CH = 8;
  const int lane = threadIdx.x & 63;
  const int N = a.N, n = a.n, nx = a.nx, nr = a.nr;
  const int g = PACKED ? mv_group(n, nr) : 64;   // lanes per block row
  const int bpw = PACKED ? 64 / g : 1;
  const int k0 = __builtin_amdgcn_readfirstlane((bid * 4 + int(threadIdx.x >> 6)) * bpw);
  if (k0 > N) return;
  const int sub = PACKED ? lane / g : 0, r = PACKED ? lane - sub * g : lane;
  const int k = k0 + sub;                       // this lane's block row
  const bool live = sub < bpw && k <= N;
  const int kc = live ? k : k0;                 // clamped: every address stays valid
  const int base = PACKED ? sub * g : 0;        // lane holding column 0 of this sub-row
  auto bcast = [&](double v, int c) -> double {
    if constexpr (PACKED) return __shfl(v, base + c, 64);
    else return wave::rl(v, c);
  };
  const double* xk = a.x + size_t(kc) * n;
  const double* xb = a.x + size_t(N + 1) * n + a.qx;
  const double xa = r < n ? mag<ABS>(xk[r < n ? r : 0]) : 0.0;   // x_k, one entry per lane
  if (k0 < N) {
    const bool row = live && k < N && r < n;
    const int kr = row ? k : k0;
    // blocks from the wave's first block row on (one base for the wave),
    // this lane's row through a 32-bit offset
    const TA* Dk = a.D + size_t(k0) * n * n;
    const TA* Ek = a.E + size_t(k0) * n * n;
    const TA* Bk = a.B + size_t(k0) * n * nx;
    const int rr = r < n ? r : 0;
    const unsigned odk = unsigned(kr - k0) * unsigned(n * n) + rr, obk = unsigned(kr - k0) * unsigned(n * nx) + rr;
    const double xn = mag<ABS>((row ? xk : a.x + size_t(k0) * n)[n + rr]);   // x_{k+1}
    // the border partial C_k x_k of the same columns rides along (lanes
    // (sub-row, i) with i < nr; a sub-row k = N has only this part)
    const bool crow = live && r < nr;
    const TA* Ck = a.C + size_t(k0) * nr * n;
    const unsigned ock = unsigned(kc - k0) * unsigned(nr * n) + (crow ? r : 0);
    // the first columns of B_k xb and this row's entry of b, in flight over
    // the column loop below
    constexpr int CB = 4;
    TA bb0[CB];
    double w0[CB];
#pragma unroll
    for (int i = 0; i < CB; ++i) {
      const int cc = i < nx ? i : 0;
      bb0[i] = nx > 0 ? gload(Bk, obk + unsigned(cc * n)) : TA(0);
      w0[i] = nx > 0 ? mag<ABS>(xb[cc]) : 0.0;
    }
    const size_t out = size_t(kr) * n + rr;
    const double bval = a.b ? a.b[out] : 0.0;
    double acc = 0, accc = 0;
    for (int c0 = 0; c0 < n; c0 += CH) {
      TA d[CH], e[CH], cv[CH];
#pragma unroll
      for (int i = 0; i < CH; ++i) {
        const int cc = c0 + i < n ? c0 + i : 0;
        d[i] = gload(Dk, odk + unsigned(cc * n));
        e[i] = gload(Ek, odk + unsigned(cc * n));
        cv[i] = nr > 0 ? gload(Ck, ock + unsigned(cc * nr)) : TA(0);
      }
#pragma unroll
      for (int i = 0; i < CH; ++i) {
        const int cc = c0 + i < n ? c0 + i : 0;
        const double u = bcast(xa, cc), v = bcast(xn, cc);
        acc += keep_if(mag<ABS>(double(d[i])) * u + mag<ABS>(double(e[i])) * v, c0 + i < n);
        accc += keep_if(mag<ABS>(double(cv[i])) * u, c0 + i < n);
      }
    }
    if (nr > 0 && crow) a.Cpart[size_t(k) * nr + r] = accc;
#pragma unroll
    for (int i = 0; i < CB; ++i) acc += keep_if(mag<ABS>(double(bb0[i])) * w0[i], i < nx);
    for (int c0 = CB; c0 < nx; c0 += CH) {
      TA bb[CH];
      double w[CH];
#pragma unroll
      for (int i = 0; i < CH; ++i) {
        const int cc = c0 + i < nx ? c0 + i : 0;
        bb[i] = gload(Bk, obk + unsigned(cc * n));
        w[i] = mag<ABS>(xb[cc]);
      }
#pragma unroll
      for (int i = 0; i < CH; ++i) acc += keep_if(mag<ABS>(double(bb[i])) * w[i], c0 + i < nx);
    }
    if (row) a.out[out] = ABS ? acc + fabs(bval) : (a.b ? bval - acc : acc);
  }
  // border partials C_k x_k of a wave that starts at block row N (the
  // others formed them above): lanes (sub-row, i) with i < nr
  if (nr > 0 && k0 >= N) {
    const int i = r;
    const bool crow = live && i < nr;
    const TA* Ck = a.C + size_t(k0) * nr * n;
    const int ic = crow ? i : 0;
    const unsigned ock = unsigned(kc - k0) * unsigned(nr * n) + ic;
    double acc = 0;
    for (int c0 = 0; c0 < n; c0 += CH) {
      TA cv[CH];
#pragma unroll
      for (int t = 0; t < CH; ++t) {
        const int cc = c0 + t < n ? c0 + t : 0;
        cv[t] = gload(Ck, ock + unsigned(cc * nr));
      }
#pragma unroll
      for (int t = 0; t < CH; ++t) {
        const int cc = c0 + t < n ? c0 + t : 0;
        acc += keep_if(mag<ABS>(double(cv[t])) * bcast(xa, cc), c0 + t < n);
      }
    }
    if (crow) a.Cpart[size_t(k) * nr + i] = acc;
  }
}

// r_q = b_q - A x_q for the nrhs columns q (x_q = x + q ldx, b_q, out_q
// likewise; border partials of column q at Cpart + q (N+1) nr): the rows of
// mv_rows_run with each block's loads shared by QB columns at a time
template <typename TA, bool PACKED, int QB>
__device__ __forceinline__ void mv_rows_multi_run(MvArgs<TA> const& a, int bid) {
  constexpr int CH = 8;
  const int lane = threadIdx.x & 63;
  const int N = a.N, n = a.n, nx = a.nx, nr = a.nr;
  const int g = PACKED ? mv_group(n, nr) : 64;
  const int bpw = PACKED ? 64 / g : 1;
  const int k0 = __builtin_amdgcn_readfirstlane((bid * 4 + int(threadIdx.x >> 6)) * bpw);
  if (k0 > N) return;
  const int sub = PACKED ? lane / g : 0, r = PACKED ? lane - sub * g : lane;
  const int k = k0 + sub;
  const bool live = sub < bpw && k <= N;
  const int kc = live ? k : k0;
  const int base = PACKED ? sub * g : 0;
  auto bcast = [&](double v, int c) -> double {
    if constexpr (PACKED) return __shfl(v, base + c, 64);
    else return wave::rl(v, c);
  };
  const bool row = live && k < N && r < n;
  const int kr = row ? k : k0;
  const int rr = r < n ? r : 0;
  const TA* Dk = a.D + size_t(k0) * n * n;
  const TA* Ek = a.E + size_t(k0) * n * n;
  const TA* Bk = a.B + size_t(k0) * n * nx;
  const TA* Ck = a.C + size_t(k0) * nr * n;
  const unsigned odk = unsigned(kr - k0) * unsigned(n * n) + rr, obk = unsigned(kr - k0) * unsigned(n * nx) + rr;
  const bool crow = live && r < nr;
  const unsigned ock = unsigned(kc - k0) * unsigned(nr * n) + (crow ? r : 0);
  const size_t out = size_t(kr) * n + rr;
  for (int q0 = 0; q0 < a.nrhs; q0 += QB) {
    double xa[QB], xn[QB], acc[QB], accc[QB];
#pragma unroll
    for (int b = 0; b < QB; ++b) {
      const int q = q0 + b < a.nrhs ? q0 + b : a.nrhs - 1;
      const double* x = a.x + size_t(q) * a.ldx;
      xa[b] = r < n ? x[size_t(kc) * n + rr] : 0.0;
      xn[b] = k0 < N ? (row ? x + size_t(k) * n : x + size_t(k0) * n)[n + rr] : 0.0;
      acc[b] = 0.0;
      accc[b] = 0.0;
    }
    if (k0 < N) {
      for (int c0 = 0; c0 < n; c0 += CH) {
        TA d[CH], e[CH], cv[CH];
#pragma unroll
        for (int i = 0; i < CH; ++i) {
          const int cc = c0 + i < n ? c0 + i : 0;
          d[i] = gload(Dk, odk + unsigned(cc * n));
          e[i] = gload(Ek, odk + unsigned(cc * n));
          cv[i] = nr > 0 ? gload(Ck, ock + unsigned(cc * nr)) : TA(0);
        }
#pragma unroll
        for (int i = 0; i < CH; ++i) {
          const int cc = c0 + i < n ? c0 + i : 0;
#pragma unroll
          for (int b = 0; b < QB; ++b) {
            const double u = bcast(xa[b], cc), v = bcast(xn[b], cc);
            acc[b] += keep_if(double(d[i]) * u + double(e[i]) * v, c0 + i < n);
            accc[b] += keep_if(double(cv[i]) * u, c0 + i < n);
          }
        }
      }
      for (int c0 = 0; c0 < nx; c0 += CH) {
        TA bb[CH];
#pragma unroll
        for (int i = 0; i < CH; ++i) bb[i] = gload(Bk, obk + unsigned((c0 + i < nx ? c0 + i : 0) * n));
#pragma unroll
        for (int b = 0; b < QB; ++b) {
          const int q = q0 + b < a.nrhs ? q0 + b : a.nrhs - 1;
          const double* xb = a.x + size_t(q) * a.ldx + size_t(N + 1) * n + a.qx;
#pragma unroll
          for (int i = 0; i < CH; ++i) acc[b] += keep_if(double(bb[i]) * xb[c0 + i < nx ? c0 + i : 0], c0 + i < nx);
        }
      }
    } else if (nr > 0) {
      // a wave that starts at block row N: the border partials C_N x_N only
      for (int c0 = 0; c0 < n; c0 += CH) {
        TA cv[CH];
#pragma unroll
        for (int i = 0; i < CH; ++i) cv[i] = gload(Ck, ock + unsigned((c0 + i < n ? c0 + i : 0) * nr));
#pragma unroll
        for (int i = 0; i < CH; ++i) {
          const int cc = c0 + i < n ? c0 + i : 0;
#pragma unroll
          for (int b = 0; b < QB; ++b) accc[b] += keep_if(double(cv[i]) * bcast(xa[b], cc), c0 + i < n);
        }
      }
    }
#pragma unroll
    for (int b = 0; b < QB; ++b) {
      const int q = q0 + b;
      if (q < a.nrhs) {
        if (nr > 0 && crow) a.Cpart[size_t(q) * (N + 1) * nr + size_t(k) * nr + r] = accc[b];
        if (row) {
          const double bv = a.b ? a.b[size_t(q) * a.ldb + out] : 0.0;
          a.out[size_t(q) * a.ldo + out] = a.b ? bv - acc[b] : acc[b];
        }
      }
    }
  }
}

// Block row k < N on one wave when n + nr <= 64: its n rows on lanes [0, n)
// and its nr border partials C_k x_k on lanes [n, n + nr).  The border lanes
// read C_k through the same load instructions as the row lanes read D_k (one
// pointer and column stride per lane), so a column costs two loads instead of
// three; their E_k products are multiplied by zero.  Sums and their order are
// those of mv_rows_run.  (The same split for the packed layout, n + nr lanes
// per block row, was no faster there: fewer block rows per wave and more
// registers ate the saved loads.)
template <typename TA, bool ABS>
__device__ __forceinline__ void mv_rows_wide(MvArgs<TA> const& a, int k) {
  constexpr int CH = 8, CB = 2;
  const int lane = threadIdx.x & 63;
  const int n = a.n, nx = a.nx, nr = a.nr;
  const bool row = lane < n, brow = !row && lane < n + nr;
  const int rr = row ? lane : 0;
  const double* xk = a.x + size_t(k) * n;
  const double* xb = a.x + size_t(a.N + 1) * n + a.qx;
  const double xa = mag<ABS>(xk[rr]);       // x_k (lane c holds column c for c < n)
  const double xn = mag<ABS>(xk[n + rr]);   // x_{k+1}
  const TA* P = brow ? a.C + size_t(k) * nr * n + (lane - n) : a.D + size_t(k) * n * n + rr;
  const unsigned ps = brow ? unsigned(nr) : unsigned(n);   // column stride of this lane's block
  const TA* Ek = a.E + size_t(k) * n * n;
  const TA* Bk = a.B + size_t(k) * n * nx;
  TA bb0[CB];
  double w0[CB];
#pragma unroll
  for (int i = 0; i < CB; ++i) {
    const int cc = i < nx ? i : 0;
    bb0[i] = nx > 0 ? gload(Bk, unsigned(rr + cc * n)) : TA(0);
    w0[i] = nx > 0 ? mag<ABS>(xb[cc]) : 0.0;
  }
  const size_t out = size_t(k) * n + rr;
  const double bval = a.b ? a.b[out] : 0.0;
  double acc = 0;
  for (int c0 = 0; c0 < n; c0 += CH) {
    TA d[CH], e[CH];
#pragma unroll
    for (int i = 0; i < CH; ++i) {
      const int cc = c0 + i < n ? c0 + i : 0;
      d[i] = gload(P, unsigned(cc) * ps);
      e[i] = gload(Ek, unsigned(rr + cc * n));
    }
#pragma unroll
    for (int i = 0; i < CH; ++i) {
      const int cc = c0 + i < n ? c0 + i : 0;
      const double u = wave::rl(xa, cc), v = brow ? 0.0 : wave::rl(xn, cc);
      acc += keep_if(mag<ABS>(double(d[i])) * u + mag<ABS>(double(e[i])) * v, c0 + i < n);
    }
  }
  if (brow) a.Cpart[size_t(k) * nr + (lane - n)] = acc;
#pragma unroll
  for (int i = 0; i < CB; ++i) acc += keep_if(mag<ABS>(double(bb0[i])) * w0[i], i < nx);
  for (int c0 = CB; c0 < nx; c0 += CH) {
    TA bb[CH];
    double w[CH];
#pragma unroll
    for (int i = 0; i < CH; ++i) {
      const int cc = c0 + i < nx ? c0 + i : 0;
      bb[i] = gload(Bk, unsigned(rr + cc * n));
      w[i] = mag<ABS>(xb[cc]);
    }
#pragma unroll
    for (int i = 0; i < CH; ++i) acc += keep_if(mag<ABS>(double(bb[i])) * w[i], c0 + i < nx);
  }
  if (row) a.out[out] = ABS ? acc + fabs(bval) : (a.b ? bval - acc : acc);
}

template <typename TA, bool ABS = false>
__device__ __forceinline__ void mv_rows_body(MvArgs<TA> const& a, int bid) {
  if (mv_rows_per_wave(a.n, a.nr) > 1) {
    mv_rows_run<TA, true, ABS>(a, bid);
  } else if (a.nr > 0 && a.n + a.nr <= 64) {
    const int k = __builtin_amdgcn_readfirstlane(bid * 4 + int(threadIdx.x >> 6));
    if (k < a.N) mv_rows_wide<TA, ABS>(a, k);
    else mv_rows_run<TA, false, ABS>(a, bid);   // block row N: its border partial (nothing past it)
  } else {
    mv_rows_run<TA, false, ABS>(a, bid);
  }
}

template <typename TA, bool ABS>
__global__ void __launch_bounds__(256) mv_rows_kernel(MvArgs<TA> a) {
  mv_rows_body<TA, ABS>(a, blockIdx.x);
}

template <typename TA>
__global__ void __launch_bounds__(256) mv_rows_multi_kernel(MvArgs<TA> a) {
  if (mv_rows_per_wave(a.n, a.nr) > 1) mv_rows_multi_run<TA, true, 4>(a, blockIdx.x);
  else mv_rows_multi_run<TA, false, 4>(a, blockIdx.x);
}


template <typename TA, bool ABS = false>
__device__ __forceinline__ void mv_tail_body(MvArgs<TA> const& a, const double* __restrict__ Csum) {
  const int N = a.N, n = a.n, qr = a.qr, qx = a.qx, nr = a.nr, nx = a.nx;
  const int m = n + qr;
  const int Nc = 2 * n + qx + nx;
  const double* x0 = a.x;
  const double* xN = a.x + size_t(N) * n;
  const double* xq = a.x + size_t(N + 1) * n;
  const double* xb = xq + qx;
  // x entry of boundary column c (x_0, x_N, q, xb in turn; c < Nc)
  auto xcol = [&](int c) -> double {
    return c < n ? x0[c] : c < 2 * n ? xN[c - n] : c < 2 * n + qx ? xq[c - 2 * n] : xb[c - 2 * n - qx];
  };
  for (int i = threadIdx.x; i < m + nr; i += blockDim.x) {
    double acc = 0;
    size_t row;
    if (i < m) {
      // eight columns' loads in flight at a time (clamped columns, the
      // products of the extra ones dropped); the sum keeps column order
      constexpr int U = 8;
      for (int c0 = 0; c0 < Nc; c0 += U) {
        double hv[U], xv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int c = c0 + u < Nc ? c0 + u : Nc - 1;
          hv[u] = double(a.H[i + size_t(c) * m]);
          xv[u] = xcol(c);
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
          if (c0 + u < Nc) acc += mag<ABS>(hv[u]) * mag<ABS>(xv[u]);
      }
      row = size_t(N) * n + i;
    } else {
      const int r = i - m;
      acc = Csum[r];
      for (int c = 0; c < qx; ++c) acc += mag<ABS>(double(a.Cq[r + c * nr])) * mag<ABS>(xq[c]);
      for (int c = 0; c < nx; ++c) acc += mag<ABS>(double(a.F[r + c * nr])) * mag<ABS>(xb[c]);
      row = size_t(N + 1) * n + qr + r;
    }
    a.out[row] = ABS ? acc + (a.b ? fabs(a.b[row]) : 0.0) : (a.b ? a.b[row] - acc : acc);
  }
}

template <typename TA, bool ABS>
__global__ void __launch_bounds__(256) mv_tail_kernel(MvArgs<TA> a, const double* __restrict__ Csum) {
  mv_tail_body<TA, ABS>(a, Csum);
}

template <typename TA, bool ABS>
__global__ void __launch_bounds__(256) mv_rows_batched(BatchTable<MvArgs<TA>> t) {
  const int e = t.entry[blockIdx.x];
  if (skipped(t.args[e].skip)) return;
  mv_rows_body<TA, ABS>(t.args[e], int(blockIdx.x) - t.first[e]);
}

// one workgroup per system; Csum[s] = the system's summed border partials
template <typename TA, bool ABS>
__global__ void __launch_bounds__(256) mv_tail_batched(const MvArgs<TA>* args, const double* const* Csum) {
  if (skipped(args[blockIdx.x].skip)) return;
  mv_tail_body<TA, ABS>(args[blockIdx.x], Csum[blockIdx.x]);
}

template <typename TA>
hipError_t launch_mv_rows_batched(BatchTable<MvArgs<TA>> const& t, int nblocks, hipStream_t st, bool absval) {
  if (absval) hipLaunchKernelGGL((mv_rows_batched<TA, true>), dim3(nblocks), dim3(256), 0, st, t);
  else hipLaunchKernelGGL((mv_rows_batched<TA, false>), dim3(nblocks), dim3(256), 0, st, t);
  return hipGetLastError();
}

template <typename TA>
hipError_t launch_mv_tail_batched(const MvArgs<TA>* args, const double* const* Csum, int nsys, hipStream_t st,
                                  bool absval) {
  if (absval) hipLaunchKernelGGL((mv_tail_batched<TA, true>), dim3(nsys), dim3(256), 0, st, args, Csum);
  else hipLaunchKernelGGL((mv_tail_batched<TA, false>), dim3(nsys), dim3(256), 0, st, args, Csum);
  return hipGetLastError();
}

template hipError_t launch_mv_rows_batched<double>(BatchTable<MvArgs<double>> const&, int, hipStream_t, bool);
template hipError_t launch_mv_tail_batched<double>(const MvArgs<double>*, const double* const*, int, hipStream_t,
                                                   bool);

// sum of squares of v[0..len) -> tmp[blockIdx] (pass 1) ; out = sqrt(sum) (pass 2)
__global__ void __launch_bounds__(256) sumsq_partial_kernel(const double* __restrict__ v, int64_t len,
                                                            double* __restrict__ tmp) {
  __shared__ double red[256];
  double s = 0;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < len; i += int64_t(gridDim.x) * blockDim.x)
    s += v[i] * v[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w >= 1; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) tmp[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(256) sumsq_final_kernel(const double* __restrict__ tmp, int nb,
                                                          double* __restrict__ out) {
  __shared__ double red[256];
  double s = 0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) s += tmp[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w >= 1; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sqrt(red[0]);
}

template <typename TA>
hipError_t launch_mv(MvArgs<TA> const& a, double* Csum, double* tmp, int max_chunks, hipStream_t st, bool absval) {
  const int rows_per_block = 4 * mv_rows_per_wave(a.n, a.nr);
  const dim3 grid((a.N + 1 + rows_per_block - 1) / rows_per_block);
  if (a.nrhs > 1 && !absval) {
    // several columns: one pass over the blocks for up to four of them, then
    // per column the border sums and the boundary / border rows (Cpart and
    // Csum hold nrhs column slices)
    hipLaunchKernelGGL((mv_rows_multi_kernel<TA>), grid, dim3(256), 0, st, a);
    for (int q = 0; q < a.nrhs; ++q) {
      MvArgs<TA> c = a;
      c.nrhs = 1;
      c.x = a.x + size_t(q) * a.ldx;
      c.b = a.b ? a.b + size_t(q) * a.ldb : nullptr;
      c.out = a.out + size_t(q) * a.ldo;
      c.Cpart = a.Cpart + size_t(q) * (a.N + 1) * a.nr;
      double* cs = Csum + size_t(q) * a.nr;
      if (a.nr > 0) {
        launch_convert<double, double>(nullptr, cs, a.nr, st);
        launch_reduce<double, double>(c.Cpart, 0, a.N + 1, a.nr, tmp, max_chunks, cs, a.nr, a.nr, st);
      }
      hipLaunchKernelGGL((mv_tail_kernel<TA, false>), dim3(1), dim3(256), 0, st, c, cs);
    }
    return hipGetLastError();
  }
  if (absval) hipLaunchKernelGGL((mv_rows_kernel<TA, true>), grid, dim3(256), 0, st, a);
  else hipLaunchKernelGGL((mv_rows_kernel<TA, false>), grid, dim3(256), 0, st, a);
  if (a.nr > 0) {
    launch_convert<double, double>(nullptr, Csum, a.nr, st);
    launch_reduce<double, double>(a.Cpart, 0, a.N + 1, a.nr, tmp, max_chunks, Csum, a.nr, a.nr, st);
  }
  if (absval) hipLaunchKernelGGL((mv_tail_kernel<TA, true>), dim3(1), dim3(256), 0, st, a, Csum);
  else hipLaunchKernelGGL((mv_tail_kernel<TA, false>), dim3(1), dim3(256), 0, st, a, Csum);
  return hipGetLastError();
}

// max_i |r_i| / den_i over i < len (0 / 0 counts as 0, r / 0 as +inf):
// pass 1 per chunk into tmp, pass 2 into *out
__global__ void __launch_bounds__(256) max_ratio_partial_kernel(const double* __restrict__ r,
                                                                const double* __restrict__ den, int64_t len,
                                                                double* __restrict__ tmp) {
  __shared__ double red[256];
  double s = 0;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < len; i += int64_t(gridDim.x) * blockDim.x) {
    const double ar = fabs(r[i]), d = den[i];
    const double q = ar == 0.0 ? 0.0 : (d > 0.0 ? ar / d : INFINITY);
    s = q > s || q != q ? q : s;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w >= 1; w >>= 1) {
    if (threadIdx.x < w) {
      const double o = red[threadIdx.x + w];
      if (o > red[threadIdx.x] || o != o) red[threadIdx.x] = o;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) tmp[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(256) max_final_kernel(const double* __restrict__ tmp, int nb,
                                                        double* __restrict__ out) {
  __shared__ double red[256];
  double s = 0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) s = tmp[i] > s || tmp[i] != tmp[i] ? tmp[i] : s;
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w >= 1; w >>= 1) {
    if (threadIdx.x < w) {
      const double o = red[threadIdx.x + w];
      if (o > red[threadIdx.x] || o != o) red[threadIdx.x] = o;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

hipError_t launch_max_ratio(const double* r, const double* den, int64_t len, double* tmp, int max_chunks, double* out,
                            hipStream_t st) {
  int64_t nb64 = (len + 255) / 256;
  if (nb64 < 1) nb64 = 1;
  if (nb64 > max_chunks) nb64 = max_chunks;
  const int nb = int(nb64);
  hipLaunchKernelGGL(max_ratio_partial_kernel, dim3(nb), dim3(256), 0, st, r, den, len, tmp);
  hipLaunchKernelGGL(max_final_kernel, dim3(1), dim3(256), 0, st, tmp, nb, out);
  return hipGetLastError();
}

hipError_t launch_nrm2(const double* v, int64_t len, double* tmp, int max_chunks, double* out, hipStream_t st) {
  int64_t nb64 = (len + 255) / 256;
  if (nb64 < 1) nb64 = 1;
  if (nb64 > max_chunks) nb64 = max_chunks;
  const int nb = int(nb64);
  hipLaunchKernelGGL(sumsq_partial_kernel, dim3(nb), dim3(256), 0, st, v, len, tmp);
  hipLaunchKernelGGL(sumsq_final_kernel, dim3(1), dim3(256), 0, st, tmp, nb, out);
  return hipGetLastError();
}

template hipError_t launch_mv<double>(MvArgs<double> const&, double*, double*, int, hipStream_t, bool);
template hipError_t launch_mv<float>(MvArgs<float> const&, double*, double*, int, hipStream_t, bool);

}  // namespace bcr
