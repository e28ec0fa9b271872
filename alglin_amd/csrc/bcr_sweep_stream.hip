// Solve sweeps for LU-eliminated pairs with n <= 64, one wavefront per pair
// (four per workgroup), the factor blocks streamed from HBM column by column
// instead of held in registers or staged in LDS.
//
//   forward  (cc:1615-1719): lane l owns rows l and l + 64 of W = [x_jp; x_j]
//            (2n <= 128 rows).  The row interchanges are composed into a
//            source row per lane and applied while gathering W from x; then
//            W <- [L 0; M I]^{-1} W with x_c broadcast by v_readlane (row c < n
//            <= 64 sits in lane c); x_j <- W[0:n], x_jp <- W[n:2n];
//            Xpart[j] = -C_j' x_j (wave sums).
//   backward (cc:1805-1949): lane r < n owns row r:
//            s = x_j - D_j x_jp - E_j x_jpp - B_j xb (the vectors broadcast from
//            the lanes that hold them), then x_j = U^{-1} s bottom up.
#include "bcr_launch.hh"
#include "bcr_batch_kernels.hh"

namespace bcr {
namespace {

__device__ __forceinline__ double rlane(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane(int(b & 0xffffffff), l);
  const int hi = __builtin_amdgcn_readlane(int(b >> 32), l);
  return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// columns loaded per chunk: one memory latency per chunk
constexpr int kChunk = 16;

// QB right-hand sides per pass: every factor chunk is loaded once for all
// of them (QB = 1 for one right-hand side; the last pass of several masks
// the columns past nrhs)
template <typename TW, typename TX, int QB>
__device__ __forceinline__ void forward_lu_stream_body(SolveArgs<TW, TX> a, int bid) {
  const int lane = threadIdx.x & 63;
  const int p = bid * 4 + __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));   // wave-uniform
  if (p >= a.npairs) return;   // uniform over the wavefront
  const int n = a.n, n2 = 2 * n, nr = a.nr, nrhs = a.nrhs;
  int j, jp, jpp;
  pair_of_uniform(a, p, j, jp, jpp);
  const TW* Tj = a.Tm + size_t(j) * tsize(n);
  const int r0 = lane, r1 = lane + 64;
  const bool v0 = r0 < n2, v1 = r1 < n2;
  // factor columns kChunk at a time, from clamped addresses: one memory
  // latency per chunk (rows past 2n hold garbage that is never used); the
  // first chunk is in flight while the interchanges are composed and x is
  // gathered
  const int r0c = v0 ? r0 : 0, r1c = v1 ? r1 : 0;
  TW t0[kChunk], t1[kChunk];
  auto load_chunk = [&](int c0) {
#pragma unroll
    for (int i = 0; i < kChunk; ++i) {
      const unsigned cc = unsigned(c0 + i < n ? c0 + i : 0) * n2;
      t0[i] = gload(Tj, r0c + cc);
      t1[i] = gload(Tj, r1c + cc);
    }
  };
  load_chunk(0);
  // source rows after the interchanges (laswp over rows 1..n, 1-based ipiv)
  int src0 = r0, src1 = r1;
  const int pvl = lane < n ? a.perm[size_t(j) * n + lane] - 1 : 0;
  for (int c = 0; c < n; ++c) {
    const int pv = __builtin_amdgcn_readlane(pvl, c);
    if (pv == c) continue;
    const int sc = __builtin_amdgcn_readlane(src0, c);
    const int sp = pv < 64 ? __builtin_amdgcn_readlane(src0, pv) : __builtin_amdgcn_readlane(src1, pv - 64);
    if (lane == c) src0 = sp;
    if (pv < 64) { if (lane == pv) src0 = sc; }
    else if (lane == pv - 64) src1 = sc;
  }
  auto in_row = [&](int r) { return r < n ? size_t(jp) * n + r : size_t(j) * n + (r - n); };
  auto out_row = [&](int r) { return r < n ? size_t(j) * n + r : size_t(jp) * n + (r - n); };
  const TW* Cj = a.C + size_t(j) * nr * n;
  for (int q0 = 0; q0 < nrhs; q0 += QB) {
    TX x0[QB], x1[QB];
#pragma unroll
    for (int b = 0; b < QB; ++b) {
      const TX* x = a.x + size_t(q0 + b < nrhs ? q0 + b : nrhs - 1) * a.ldx;
      x0[b] = v0 ? x[in_row(src0)] : TX(0);
      x1[b] = v1 ? x[in_row(src1)] : TX(0);
    }
    for (int c0 = 0; c0 < n; c0 += kChunk) {
      if (c0 > 0 || q0 > 0) load_chunk(c0);
#pragma unroll
      for (int i = 0; i < kChunk; ++i) {
        const int c = c0 + i;
        if (c < n) {
#pragma unroll
          for (int b = 0; b < QB; ++b) {
            const TX xc = TX(rlane(double(x0[b]), c));
            if (r0 > c) x0[b] -= TX(t0[i]) * xc;
            x1[b] -= TX(t1[i]) * xc;
          }
        }
      }
    }
#pragma unroll
    for (int b = 0; b < QB; ++b) {
      const int q = q0 + b;
      if (q < nrhs) {
        TX* x = a.x + size_t(q) * a.ldx;
        if (v0) x[out_row(r0)] = x0[b];
        if (v1) x[out_row(r1)] = x1[b];
        for (int i = 0; i < nr; ++i) {
          const double s = wsum(lane < n ? double(Cj[i + size_t(lane) * nr]) * double(x0[b]) : 0.0);
          if (lane == 0) a.Xpart[size_t(j) * nr * nrhs + i + size_t(q) * nr] = TX(-s);
        }
      }
    }
  }
}

template <typename TW, typename TX, int QB>
__device__ __forceinline__ void backward_lu_stream_body(SolveArgs<TW, TX> a, int bid) {
  const int lane = threadIdx.x & 63;
  const int p = bid * 4 + __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));   // wave-uniform
  if (p >= a.npairs) return;
  const int n = a.n, n2 = 2 * n, nx = a.nx, nrhs = a.nrhs;
  int j, jp, jpp;
  pair_of_uniform(a, p, j, jp, jpp);
  const size_t nn = size_t(n) * n;
  const TW* Dj = a.D + size_t(j) * nn;
  const TW* Ej = a.E + size_t(j) * nn;
  const TW* Bj = a.B + size_t(j) * n * nx;
  const TW* Tj = a.Tm + size_t(j) * tsize(n);
  const bool row = lane < n;
  const size_t xbrow = size_t(a.N + 1) * n + a.qx;   // cc:1812
  const int rc = row ? lane : 0;   // lanes past n read row 0 and never store
  // the reciprocal of this lane's diagonal of U, once per pair: the bottom-up
  // recurrence multiplies instead of dividing (x_c = s_c * (1/u_cc))
  const TX dinv = row ? TX(1) / TX(Tj[rc + size_t(rc) * n2]) : TX(1);
  // QRP: y_i <-> y_{s_i} for i = 0..n-1 with s_i > i (cc:1851-1856), composed
  // into the lane whose value ends in row `lane`
  int src = lane;
  if (a.exch) {
    const int sl = row ? a.perm[size_t(j) * n + lane] : 0;
    for (int i = 0; i < n; ++i) {
      const int s = __builtin_amdgcn_readlane(sl, i);
      if (s > i && s < n) {
        const int si = __builtin_amdgcn_readlane(src, i), ss = __builtin_amdgcn_readlane(src, s);
        src = lane == i ? ss : lane == s ? si : src;
      }
    }
  }
  if constexpr (QB == 1) {   // one column: the per-column loop as tuned
    for (int q = 0; q < nrhs; ++q) {
      TX* x = a.x + size_t(q) * a.ldx;
      const TX vjp = row ? x[size_t(jp) * n + lane] : TX(0);
      const TX vjpp = row ? x[size_t(jpp) * n + lane] : TX(0);
      TX s = row ? x[size_t(j) * n + lane] : TX(0);
      // B_j's first two columns, in flight with the first D_j, E_j chunk
      // (eight would cost this kernel two waves per SIMD)
      TW bv0[2];
  #pragma unroll
      for (int i = 0; i < 2; ++i) bv0[i] = gload(Bj, rc + unsigned(i < nx ? i : 0) * n);
      // D_j, E_j columns kChunk at a time (one memory latency per chunk)
      for (int c0 = 0; c0 < n; c0 += kChunk) {
        TW d[kChunk], e[kChunk];
  #pragma unroll
        for (int i = 0; i < kChunk; ++i) {
          const unsigned cc = unsigned(c0 + i < n ? c0 + i : 0) * n;
          d[i] = gload(Dj, rc + cc);
          e[i] = gload(Ej, rc + cc);
        }
  #pragma unroll
        for (int i = 0; i < kChunk; ++i) {
          const int c = c0 + i;
          if (c < n) s -= TX(d[i]) * TX(rlane(double(vjp), c)) + TX(e[i]) * TX(rlane(double(vjpp), c));
        }
      }
      // B_j xb: the two columns above, then eight columns' loads in flight
      // at a time
  #pragma unroll
      for (int i = 0; i < 2; ++i)
        if (i < nx) s -= TX(keep_if(bv0[i], row)) * x[xbrow + i];
      for (int c0 = 2; c0 < nx; c0 += 8) {
        TW bv[8];
        TX xv[8];
  #pragma unroll
        for (int i = 0; i < 8; ++i) {
          const int cc = c0 + i < nx ? c0 + i : 0;
          bv[i] = gload(Bj, rc + unsigned(cc) * n);
          xv[i] = x[xbrow + cc];
        }
  #pragma unroll
        for (int i = 0; i < 8; ++i)
          if (c0 + i < nx) s -= TX(keep_if(bv[i], row)) * xv[i];
      }
      // U^{-1}, bottom up: lane c owns the diagonal of column c; columns kChunk
      // at a time from the top of the remaining range down
      for (int c1 = n - 1; c1 >= 0; c1 -= kChunk) {
        TW u[kChunk];
  #pragma unroll
        for (int i = 0; i < kChunk; ++i) {
          const int c = c1 - i;
          u[i] = gload(Tj, rc + unsigned(c >= 0 ? c : 0) * n2);
        }
  #pragma unroll
        for (int i = 0; i < kChunk; ++i) {
          const int c = c1 - i;
          if (c >= 0) {
            const TX xc = TX(rlane(double(s * dinv), c));
            if (lane < c) s -= TX(u[i]) * xc;
            else if (lane == c) s = xc;
          }
        }
      }
      if (a.exch) s = __shfl(s, src);
      if (row) x[size_t(j) * n + lane] = s;
    }
  } else {
    for (int q0 = 0; q0 < nrhs; q0 += QB) {
      TX vjp[QB], vjpp[QB], s[QB];
      const TX* xq[QB];
  #pragma unroll
      for (int b = 0; b < QB; ++b) {
        xq[b] = a.x + size_t(q0 + b < nrhs ? q0 + b : nrhs - 1) * a.ldx;
        vjp[b] = row ? xq[b][size_t(jp) * n + lane] : TX(0);
        vjpp[b] = row ? xq[b][size_t(jpp) * n + lane] : TX(0);
        s[b] = row ? xq[b][size_t(j) * n + lane] : TX(0);
      }
      // D_j, E_j columns kChunk at a time (one memory latency per chunk)
      for (int c0 = 0; c0 < n; c0 += kChunk) {
        TW d[kChunk], e[kChunk];
  #pragma unroll
        for (int i = 0; i < kChunk; ++i) {
          const unsigned cc = unsigned(c0 + i < n ? c0 + i : 0) * n;
          d[i] = gload(Dj, rc + cc);
          e[i] = gload(Ej, rc + cc);
        }
  #pragma unroll
        for (int i = 0; i < kChunk; ++i) {
          const int c = c0 + i;
          if (c < n) {
  #pragma unroll
            for (int b = 0; b < QB; ++b)
              s[b] -= TX(d[i]) * TX(rlane(double(vjp[b]), c)) + TX(e[i]) * TX(rlane(double(vjpp[b]), c));
          }
        }
      }
      // B_j xb, eight columns' loads in flight at a time
      for (int c0 = 0; c0 < nx; c0 += 8) {
        TW bv[8];
  #pragma unroll
        for (int i = 0; i < 8; ++i) {
          const int cc = c0 + i < nx ? c0 + i : 0;
          bv[i] = gload(Bj, rc + unsigned(cc) * n);
        }
  #pragma unroll
        for (int b = 0; b < QB; ++b) {
          TX xv[8];
  #pragma unroll
          for (int i = 0; i < 8; ++i) xv[i] = xq[b][xbrow + (c0 + i < nx ? c0 + i : 0)];
  #pragma unroll
          for (int i = 0; i < 8; ++i)
            if (c0 + i < nx) s[b] -= TX(keep_if(bv[i], row)) * xv[i];
        }
      }
      // U^{-1}, bottom up: lane c owns the diagonal of column c; columns kChunk
      // at a time from the top of the remaining range down
      for (int c1 = n - 1; c1 >= 0; c1 -= kChunk) {
        TW u[kChunk];
  #pragma unroll
        for (int i = 0; i < kChunk; ++i) {
          const int c = c1 - i;
          u[i] = gload(Tj, rc + unsigned(c >= 0 ? c : 0) * n2);
        }
  #pragma unroll
        for (int i = 0; i < kChunk; ++i) {
          const int c = c1 - i;
          if (c >= 0) {
  #pragma unroll
            for (int b = 0; b < QB; ++b) {
              const TX xc = TX(rlane(double(s[b] * dinv), c));
              if (lane < c) s[b] -= TX(u[i]) * xc;
              else if (lane == c) s[b] = xc;
            }
          }
        }
      }
  #pragma unroll
      for (int b = 0; b < QB; ++b) {
        if (a.exch) s[b] = __shfl(s[b], src);
        if (row && q0 + b < nrhs) a.x[size_t(q0 + b) * a.ldx + size_t(j) * n + lane] = s[b];
      }
    }
  }
}

// Back substitution for n <= 32 with two pairs per wavefront: lanes
// 32 h + r (r < n) own row r of pair 2 w + h, so no lane idles on the upper
// half; a vector entry of column c is broadcast per half (v_readlane of
// lanes c and 32 + c, then a select).  One right-hand side, LU exchanges
// only, pairs in pure cyclic-reduction order (j grows with the pair index:
// the wavefront's second pair is addressed from its first one's blocks).
template <typename TW, typename TX>
__device__ __forceinline__ void backward_lu_stream2_body(SolveArgs<TW, TX> a, int bid) {
  const int lane = threadIdx.x & 63, h = lane >> 5, r = lane & 31;
  const int p0 = (bid * 2 + __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6))) * 2;   // wave-uniform
  if (p0 >= a.npairs) return;
  const int p = p0 + h;
  const bool live = p < a.npairs;
  const int n = a.n, n2 = 2 * n, nx = a.nx;
  int j0, jp0, jpp0, j, jp, jpp;
  pair_index(p0, a.L, a.N, j0, jp0, jpp0);
  pair_index(live ? p : p0, a.L, a.N, j, jp, jpp);
  const unsigned dj = unsigned(j - j0);
  const size_t nn = size_t(n) * n;
  const TW* Dw = a.D + size_t(j0) * nn;
  const TW* Ew = a.E + size_t(j0) * nn;
  const TW* Bw = a.B + size_t(j0) * n * nx;
  const TW* Tw = a.Tm + size_t(j0) * tsize(n);
  const bool row = live && r < n;
  const int rc = row ? r : 0;
  const unsigned od = dj * unsigned(nn) + unsigned(rc), ob = dj * unsigned(n * nx) + unsigned(rc);
  const unsigned ot = dj * unsigned(tsize(n)) + unsigned(rc);
  auto bc = [&](TX v, int c) -> TX {
    const TX v0 = TX(rlane(double(v), c)), v1 = TX(rlane(double(v), 32 + c));
    return h ? v1 : v0;
  };
  const TX dinv = row ? TX(1) / TX(gload(Tw, ot + unsigned(rc) * n2)) : TX(1);
  const size_t xbrow = size_t(a.N + 1) * n + a.qx;   // cc:1812
  TX* const x = a.x;
  const TX vjp = row ? x[size_t(jp) * n + r] : TX(0);
  const TX vjpp = row ? x[size_t(jpp) * n + r] : TX(0);
  TX s = row ? x[size_t(j) * n + r] : TX(0);
  // B_j's first two columns, in flight with the first D_j, E_j chunk
  TW bv0[2];
#pragma unroll
  for (int i = 0; i < 2; ++i) bv0[i] = gload(Bw, ob + unsigned(i < nx ? i : 0) * n);
  // D_j x_jp + E_j x_jpp, kChunk columns per memory latency
  for (int c0 = 0; c0 < n; c0 += kChunk) {
    TW d[kChunk], e[kChunk];
#pragma unroll
    for (int i = 0; i < kChunk; ++i) {
      const unsigned cc = unsigned(c0 + i < n ? c0 + i : 0) * n;
      d[i] = gload(Dw, od + cc);
      e[i] = gload(Ew, od + cc);
    }
#pragma unroll
    for (int i = 0; i < kChunk; ++i) {
      const int c = c0 + i;
      if (c < n) s -= TX(d[i]) * bc(vjp, c) + TX(e[i]) * bc(vjpp, c);
    }
  }
  // B_j xb: the two columns above, then eight columns' loads in flight at
  // a time
#pragma unroll
  for (int i = 0; i < 2; ++i)
    if (i < nx) s -= TX(keep_if(bv0[i], row)) * x[xbrow + i];
  for (int c0 = 2; c0 < nx; c0 += 8) {
    TW bv[8];
    TX xv[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int cc = c0 + i < nx ? c0 + i : 0;
      bv[i] = gload(Bw, ob + unsigned(cc) * n);
      xv[i] = x[xbrow + cc];
    }
#pragma unroll
    for (int i = 0; i < 8; ++i)
      if (c0 + i < nx) s -= TX(keep_if(bv[i], row)) * xv[i];
  }
  // U^{-1}, bottom up: lane 32 h + c owns the diagonal of column c
  for (int c1 = n - 1; c1 >= 0; c1 -= kChunk) {
    TW u[kChunk];
#pragma unroll
    for (int i = 0; i < kChunk; ++i) {
      const int c = c1 - i;
      u[i] = gload(Tw, ot + unsigned(c >= 0 ? c : 0) * n2);
    }
#pragma unroll
    for (int i = 0; i < kChunk; ++i) {
      const int c = c1 - i;
      if (c >= 0) {
        const TX xc = bc(s * dinv, c);
        if (r < c) s -= TX(u[i]) * xc;
        else if (r == c) s = xc;
      }
    }
  }
  if (row) x[size_t(j) * n + r] = s;
}

template <typename TW, typename TX, bool FWD, int QB>
__global__ void __launch_bounds__(256) sweep_lu_stream_kernel(SolveArgs<TW, TX> a) {
  if (FWD) forward_lu_stream_body<TW, TX, QB>(a, blockIdx.x);
  else backward_lu_stream_body<TW, TX, QB>(a, blockIdx.x);
}

template <typename TW, typename TX, bool FWD>
__global__ void __launch_bounds__(256) sweep_lu_stream_batched(BatchTable<SolveArgs<TW, TX>> t) {
  const int e = t.entry[blockIdx.x];
  if (skipped(t.args[e].skip)) return;
  if (FWD) forward_lu_stream_body<TW, TX, 1>(t.args[e], int(blockIdx.x) - t.first[e]);
  else backward_lu_stream_body<TW, TX, 1>(t.args[e], int(blockIdx.x) - t.first[e]);
}

// four pairs per workgroup of two wavefronts (the batch tables' grouping)
template <typename TW, typename TX>
__global__ void __launch_bounds__(128) sweep_lu_stream2_back_batched(BatchTable<SolveArgs<TW, TX>> t) {
  const int e = t.entry[blockIdx.x];
  if (skipped(t.args[e].skip)) return;
  backward_lu_stream2_body<TW, TX>(t.args[e], int(blockIdx.x) - t.first[e]);
}

}  // namespace

template <typename TW, typename TX>
bool launch_sweep_lu_stream(bool forward, SolveArgs<TW, TX> const& a, hipStream_t st, hipError_t* err) {
  if (a.n < 1 || a.n > 64 || a.npairs < 1) return false;
  const int blocks = (a.npairs + 3) / 4;
  // several right-hand sides: four per pass over the factors
  if (a.nrhs > 1) {
    if (forward) hipLaunchKernelGGL((sweep_lu_stream_kernel<TW, TX, true, 4>), dim3(blocks), dim3(256), 0, st, a);
    else hipLaunchKernelGGL((sweep_lu_stream_kernel<TW, TX, false, 4>), dim3(blocks), dim3(256), 0, st, a);
  } else {
    if (forward) hipLaunchKernelGGL((sweep_lu_stream_kernel<TW, TX, true, 1>), dim3(blocks), dim3(256), 0, st, a);
    else hipLaunchKernelGGL((sweep_lu_stream_kernel<TW, TX, false, 1>), dim3(blocks), dim3(256), 0, st, a);
  }
  *err = hipGetLastError();
  return true;
}

template <typename TW, typename TX>
bool launch_sweep_lu_stream_batched(bool forward, BatchTable<SolveArgs<TW, TX>> const& t, int nblocks, int n,
                                    hipStream_t st, hipError_t* err) {
  if (n < 1 || n > 64) return false;
  // (the batch tables hold pure cyclic-reduction pairs, four per workgroup)
  if (forward) hipLaunchKernelGGL((sweep_lu_stream_batched<TW, TX, true>), dim3(nblocks), dim3(256), 0, st, t);
  else if (n <= 32) hipLaunchKernelGGL((sweep_lu_stream2_back_batched<TW, TX>), dim3(nblocks), dim3(128), 0, st, t);
  else hipLaunchKernelGGL((sweep_lu_stream_batched<TW, TX, false>), dim3(nblocks), dim3(256), 0, st, t);
  *err = hipGetLastError();
  return true;
}

#define BCR_INST(TW, TX)                                                                                         \
  template bool launch_sweep_lu_stream<TW, TX>(bool, SolveArgs<TW, TX> const&, hipStream_t, hipError_t*);       \
  template bool launch_sweep_lu_stream_batched<TW, TX>(bool, BatchTable<SolveArgs<TW, TX>> const&, int, int,     \
                                                       hipStream_t, hipError_t*);
BCR_INST(double, double)
BCR_INST(float, float)
BCR_INST(float, double)
#undef BCR_INST

}  // namespace bcr
