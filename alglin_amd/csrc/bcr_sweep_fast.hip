// Wave-synchronous solve sweeps for LU-eliminated pairs, n <= 32.
//
// A pair is owned by a group of G = 2*NMAX lanes of one wavefront (NMAX in
// {4, 8, 16, 32}; 64/G pairs per wavefront, four wavefronts per workgroup).
// Lane r of the group holds row r of the pair's factor blocks in registers,
// so the short recurrences (row interchanges, unit-lower and upper triangular
// solves) run with cross-lane broadcasts (readlane when the group is the whole
// wavefront, ds_bpermute otherwise) and no workgroup barrier.  Every
// right-hand side of a chunk of QB is carried in its own register.
//
//   forward  (cc:1615-1719): W = [x_jp; x_j];  W <- P W;  W <- [L 0; M I]^{-1} W
//            x_j <- W[0:n], x_jp <- W[n:2n]  (cc:1041-1042)
//            Xpart[j] = -C_j' x_j            (cc:1644-1650)
//   backward (cc:1805-1949): x_j <- U^{-1} (x_j - D_j x_jp - E_j x_jpp - B_j xb)
//            (half 0 of the group forms D_j x_jp + B_j xb, half 1 E_j x_jpp)
#include "bcr_launch.hh"

namespace bcr {
namespace {

__device__ __forceinline__ double rlane(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane(int(b & 0xffffffff), l);
  const int hi = __builtin_amdgcn_readlane(int(b >> 32), l);
  return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}
__device__ __forceinline__ float rlane(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// value of lane base+idx of the group; for G == 64 idx must be wave-uniform
template <int G, typename T>
__device__ __forceinline__ T gbcast(T v, int base, int idx) {
  if constexpr (G == 64) return rlane(v, idx);
  else return __shfl(v, base + idx, 64);
}

template <int G, typename T>
__device__ __forceinline__ T group_sum(T v) {
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <typename TW, typename TX, int NMAX, int QB>
__device__ __forceinline__ void forward_lu_fast_body(SolveArgs<TW, TX> a, int bid) {
  constexpr int G = 2 * NMAX, PG = 64 / G;
  const int lane = threadIdx.x & 63;
  const int r = lane % G, base = lane - r;
  const int p = (int(bid) * 4 + __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6))) * PG + lane / G;
  const bool live = p < a.npairs;
  const int n = a.n, n2 = 2 * n, nr = a.nr, nrhs = a.nrhs;
  const int64_t ldx = a.ldx;
  int j = 0, jp = 0, jpp = 0;
  if (live) pair_of(a, p, j, jp, jpp);
  if constexpr (G == 64) {   // one pair per wavefront: scalar block bases
    j = __builtin_amdgcn_readfirstlane(j);
    jp = __builtin_amdgcn_readfirstlane(jp);
    jpp = __builtin_amdgcn_readfirstlane(jpp);
  }
  const bool row = live && r < n2;

  // row r of the 2n x n panel [L; M] (column-major, ld 2n)
  // loads are unconditional (clamped to valid addresses, masked after) so
  // that they are all in flight together instead of one branch and one
  // memory latency each
  const TW* Tj = a.Tm + size_t(j) * tsize(n);
  const int rc = row ? r : 0;
  TW t[NMAX];   // kept in the factor precision, widened where used
#pragma unroll
  for (int c = 0; c < NMAX; ++c) {
    const TW v = gload(Tj, rc + unsigned(c < n ? c : 0) * n2);
    t[c] = keep_if(v, row && c < n);
  }

  // row interchanges (laswp over rows 1..n with 1-based ipiv, cc:939):
  // after the swaps row r holds the original row src
  int src = r;
  {
    const int pvl = (live && r < n) ? a.perm[size_t(j) * n + r] - 1 : r;
    for (int c = 0; c < n; ++c) {
      int pv, sc, sp;
      if constexpr (G == 64) {   // one pair per wavefront: c and pv are uniform
        pv = __builtin_amdgcn_readlane(pvl, c);
        sc = __builtin_amdgcn_readlane(src, c);
        sp = __builtin_amdgcn_readlane(src, pv);
      } else {
        pv = __shfl(pvl, base + c, 64);
        sc = __shfl(src, base + c, 64);
        sp = __shfl(src, base + pv, 64);
      }
      src = (r == c) ? sp : ((r == pv) ? sc : src);
    }
  }
  const size_t in_row = r < n ? size_t(jp) * n + r : size_t(j) * n + (r - n);
  const size_t out_row = r < n ? size_t(j) * n + r : size_t(jp) * n + (r - n);
  const TW* Cj = a.C + size_t(j) * nr * n;

  for (int q0 = 0; q0 < nrhs; q0 += QB) {
    TX x[QB];
#pragma unroll
    for (int q = 0; q < QB; ++q) {
      const int qq = q0 + q < nrhs ? q0 + q : 0;
      const TX v = a.x[(row ? in_row : size_t(0)) + size_t(qq) * ldx];
      x[q] = keep_if(v, row && q0 + q < nrhs);
    }
#pragma unroll
    for (int q = 0; q < QB; ++q) x[q] = __shfl(x[q], base + src, 64);
    // W <- [L 0; M I]^{-1} W : column c of the panel below the diagonal
#pragma unroll
    for (int c = 0; c < NMAX; ++c) {
      if (c < n) {
        TW tc = t[c];
        asm volatile("" : "+v"(tc));
#pragma unroll
        for (int q = 0; q < QB; ++q) {
          const TX xc = gbcast<G>(x[q], base, c);
          x[q] = (r > c) ? x[q] - TX(tc) * xc : x[q];
        }
      }
    }
#pragma unroll
    for (int q = 0; q < QB; ++q)
      if (row && q0 + q < nrhs) a.x[out_row + size_t(q0 + q) * ldx] = x[q];
    // border partial -C_j' x_j over the top n rows of the group, four border
    // rows' loads in flight at a time
    const bool crow = live && r < n;
    for (int i0 = 0; i0 < nr; i0 += 4) {
      TX ci[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int ic = i0 + i < nr ? i0 + i : 0;
        const TW v = gload(Cj, ic + unsigned(crow ? r : 0) * nr);
        ci[i] = TX(keep_if(v, crow && i0 + i < nr));
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        if (i0 + i >= nr) break;
#pragma unroll
        for (int q = 0; q < QB; ++q) {
          const TX v = group_sum<G>(ci[i] * x[q]);
          if (live && r == 0 && q0 + q < nrhs) a.Xpart[size_t(j) * nr * nrhs + i0 + i + size_t(q0 + q) * nr] = -v;
        }
      }
    }
  }
}

template <typename TW, typename TX, int NMAX, int QB>
__global__ void __launch_bounds__(256) forward_lu_fast_kernel(SolveArgs<TW, TX> a) {
  forward_lu_fast_body<TW, TX, NMAX, QB>(a, blockIdx.x);
}


template <typename TW, typename TX, int NMAX, int QB>
__device__ __forceinline__ void backward_lu_fast_body(SolveArgs<TW, TX> a, int bid) {
  constexpr int G = 2 * NMAX, PG = 64 / G;
  const int lane = threadIdx.x & 63;
  const int r = lane % G, base = lane - r;
  const int p = (int(bid) * 4 + __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6))) * PG + lane / G;
  const bool live = p < a.npairs;
  const int n = a.n, n2 = 2 * n, nx = a.nx, nrhs = a.nrhs;
  const int64_t ldx = a.ldx;
  int j = 0, jp = 0, jpp = 0;
  if (live) pair_of(a, p, j, jp, jpp);
  if constexpr (G == 64) {   // one pair per wavefront: scalar block bases
    j = __builtin_amdgcn_readfirstlane(j);
    jp = __builtin_amdgcn_readfirstlane(jp);
    jpp = __builtin_amdgcn_readfirstlane(jpp);
  }
  const int h = r / NMAX, rr = r % NMAX;
  const bool row = live && rr < n;
  const size_t nn = size_t(n) * n;

  // half 0: row rr of D_j and of U; half 1: row rr of E_j
  const TW* M = (h == 0 ? a.D : a.E) + size_t(j) * nn;
  const TW* Tj = a.Tm + size_t(j) * tsize(n);
  // loads are unconditional (clamped to valid addresses, masked after) so
  // that they are all in flight together
  // (kept in the factor precision, widened where used)
  const int rc = row ? rr : 0;
  TW m[NMAX], u[NMAX];
#pragma unroll
  for (int c = 0; c < NMAX; ++c) {
    const int cc = c < n ? c : 0;
    const TW vm = M[rc + size_t(cc) * n];
    const TW vu = Tj[rc + size_t(cc) * n2];
    m[c] = keep_if(vm, row && c < n);
    u[c] = keep_if(vu, row && h == 0 && c < n && c >= rr);
  }
  // the reciprocal of this lane's diagonal of U, formed once per pair: the
  // bottom-up recurrence then multiplies instead of dividing on its critical
  // path (x_c = s_c * (1/u_cc) rather than s_c / u_cc)
  const TW ud = Tj[rc + size_t(rc) * n2];
  const TX dinv = row && h == 0 ? TX(1) / TX(ud) : TX(1);
  const TW* Bj = a.B + size_t(j) * n * nx;
  const size_t xbrow = size_t(a.N + 1) * n + a.qx;   // cc:1812
  const size_t vin_row = size_t(h == 0 ? jp : jpp) * n + rr;

  for (int q0 = 0; q0 < nrhs; q0 += QB) {
    TX s[QB], vin[QB], xj[QB];
#pragma unroll
    for (int q = 0; q < QB; ++q) {
      const int qq = q0 + q < nrhs ? q0 + q : 0;
      const TX v = a.x[(row ? vin_row : size_t(0)) + size_t(qq) * ldx];
      const TX w = a.x[(row ? size_t(j) * n + rr : size_t(0)) + size_t(qq) * ldx];
      vin[q] = keep_if(v, row && q0 + q < nrhs);
      xj[q] = keep_if(w, row && h == 0 && q0 + q < nrhs);
      s[q] = TX(0);
    }
    // half 0: D_j x_jp, half 1: E_j x_jpp (each factor widened at its use:
    // the opaque copy keeps the widened values from being hoisted out of the
    // right-hand-side loop, which would double their registers)
#pragma unroll
    for (int c = 0; c < NMAX; ++c) {
      if (c < n) {
        TW mc = m[c];
        asm volatile("" : "+v"(mc));
#pragma unroll
        for (int q = 0; q < QB; ++q) {
          TX v;
          if constexpr (G == 64) {
            const TX v0 = rlane(vin[q], c), v1 = rlane(vin[q], NMAX + c);
            v = h == 0 ? v0 : v1;
          } else {
            v = __shfl(vin[q], base + h * NMAX + c, 64);
          }
          s[q] += TX(mc) * v;
        }
      }
    }
    // half 0: + B_j xb, four columns' loads in flight at a time
    for (int c0 = 0; c0 < nx; c0 += 4) {
      TX b[4], xb[4][QB];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int cc = c0 + c < nx ? c0 + c : 0;
        const TW v = Bj[rc + size_t(cc) * n];
        b[c] = TX(keep_if(v, row && h == 0 && c0 + c < nx));
#pragma unroll
        for (int q = 0; q < QB; ++q) {
          const int qq = q0 + q < nrhs ? q0 + q : 0;
          const TX w = a.x[xbrow + cc + size_t(qq) * ldx];
          xb[c][q] = keep_if(w, q0 + q < nrhs);
        }
      }
#pragma unroll
      for (int c = 0; c < 4; ++c)
        if (c0 + c < nx) {
#pragma unroll
          for (int q = 0; q < QB; ++q) s[q] += b[c] * xb[c][q];
        }
    }
#pragma unroll
    for (int q = 0; q < QB; ++q) {
      const TX other = __shfl(s[q], base + NMAX + rr, 64);
      s[q] = xj[q] - s[q] - other;
    }
    // U^{-1}: column c, bottom up; lane c of half 0 owns the diagonal
#pragma unroll
    for (int c = NMAX - 1; c >= 0; --c) {
      if (c < n) {
        TW uc = u[c];
        asm volatile("" : "+v"(uc));
#pragma unroll
        for (int q = 0; q < QB; ++q) {
          const TX xc = gbcast<G>(s[q] * dinv, base, c);
          s[q] = (rr < c) ? s[q] - TX(uc) * xc : ((rr == c) ? xc : s[q]);
        }
      }
    }
#pragma unroll
    for (int q = 0; q < QB; ++q)
      if (row && h == 0 && q0 + q < nrhs) a.x[size_t(j) * n + rr + size_t(q0 + q) * ldx] = s[q];
  }
}

template <typename TW, typename TX, int NMAX, int QB>
__global__ void __launch_bounds__(256) backward_lu_fast_kernel(SolveArgs<TW, TX> a) {
  backward_lu_fast_body<TW, TX, NMAX, QB>(a, blockIdx.x);
}


template <typename TW, typename TX, int NMAX, int QB, bool FWD>
__global__ void __launch_bounds__(256) sweep_lu_fast_batched(BatchTable<SolveArgs<TW, TX>> t) {
  const int e = t.entry[blockIdx.x];
  if (skipped(t.args[e].skip)) return;
  if (FWD) forward_lu_fast_body<TW, TX, NMAX, QB>(t.args[e], int(blockIdx.x) - t.first[e]);
  else backward_lu_fast_body<TW, TX, NMAX, QB>(t.args[e], int(blockIdx.x) - t.first[e]);
}

template <typename TW, typename TX, int NMAX>
hipError_t launch_sweep_nmax(bool forward, SolveArgs<TW, TX> const& a, hipStream_t st) {
  constexpr int PG = 64 / (2 * NMAX);
  const int blocks = (a.npairs + 4 * PG - 1) / (4 * PG);
  if (forward) {
    if (a.nrhs == 1) hipLaunchKernelGGL((forward_lu_fast_kernel<TW, TX, NMAX, 1>), dim3(blocks), dim3(256), 0, st, a);
    else hipLaunchKernelGGL((forward_lu_fast_kernel<TW, TX, NMAX, 4>), dim3(blocks), dim3(256), 0, st, a);
  } else {
    if (a.nrhs == 1) hipLaunchKernelGGL((backward_lu_fast_kernel<TW, TX, NMAX, 1>), dim3(blocks), dim3(256), 0, st, a);
    else hipLaunchKernelGGL((backward_lu_fast_kernel<TW, TX, NMAX, 4>), dim3(blocks), dim3(256), 0, st, a);
  }
  return hipGetLastError();
}

}  // namespace

template <typename TW, typename TX>
bool launch_sweep_lu_fast(bool forward, SolveArgs<TW, TX> const& a, hipStream_t st, hipError_t* err) {
  if (a.n < 1 || a.n > 32 || a.npairs < 1) return false;
  if (a.n <= 4) *err = launch_sweep_nmax<TW, TX, 4>(forward, a, st);
  else if (a.n <= 8) *err = launch_sweep_nmax<TW, TX, 8>(forward, a, st);
  else if (a.n <= 16) *err = launch_sweep_nmax<TW, TX, 16>(forward, a, st);
  else *err = launch_sweep_nmax<TW, TX, 32>(forward, a, st);
  return true;
}

// pairs per workgroup of the sweep kernels for block size bound n (<= 32)
int sweep_lu_fast_pairs_per_block(int n) {
  const int nmax = n <= 4 ? 4 : n <= 8 ? 8 : n <= 16 ? 16 : 32;
  return 4 * (64 / (2 * nmax));
}

template <typename TW, typename TX>
bool launch_sweep_lu_fast_batched(bool forward, BatchTable<SolveArgs<TW, TX>> const& t, int nblocks, int n, int nrhs,
                                  hipStream_t st, hipError_t* err) {
  if (n < 1 || n > 32) return false;
  auto go = [&](auto k1, auto k4) {
    if (nrhs == 1) hipLaunchKernelGGL(k1, dim3(nblocks), dim3(256), 0, st, t);
    else hipLaunchKernelGGL(k4, dim3(nblocks), dim3(256), 0, st, t);
    *err = hipGetLastError();
    return true;
  };
#define BCR_SW(NM)                                                                                                \
  return forward ? go(sweep_lu_fast_batched<TW, TX, NM, 1, true>, sweep_lu_fast_batched<TW, TX, NM, 4, true>)     \
                 : go(sweep_lu_fast_batched<TW, TX, NM, 1, false>, sweep_lu_fast_batched<TW, TX, NM, 4, false>);
  if (n <= 4) { BCR_SW(4) }
  if (n <= 8) { BCR_SW(8) }
  if (n <= 16) { BCR_SW(16) }
  BCR_SW(32)
#undef BCR_SW
}

template bool launch_sweep_lu_fast_batched<double, double>(bool, BatchTable<SolveArgs<double, double>> const&, int, int,
                                                           int, hipStream_t, hipError_t*);
template bool launch_sweep_lu_fast_batched<float, float>(bool, BatchTable<SolveArgs<float, float>> const&, int, int,
                                                         int, hipStream_t, hipError_t*);
template bool launch_sweep_lu_fast_batched<float, double>(bool, BatchTable<SolveArgs<float, double>> const&, int, int,
                                                          int, hipStream_t, hipError_t*);
template bool launch_sweep_lu_fast<double, double>(bool, SolveArgs<double, double> const&, hipStream_t, hipError_t*);
template bool launch_sweep_lu_fast<float, float>(bool, SolveArgs<float, float> const&, hipStream_t, hipError_t*);
template bool launch_sweep_lu_fast<float, double>(bool, SolveArgs<float, double> const&, hipStream_t, hipError_t*);

}  // namespace bcr
