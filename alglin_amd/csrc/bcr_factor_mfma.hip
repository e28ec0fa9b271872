// Pair elimination, LU, FP32 factors n <= 64 and FP64 factors n <= 48: the getrf + applyT split of
// the reference pair step (buildT BABD_BorderedCR.cc:856, applyT cc:917-982,
// the border update cc:1107-1154), with the right-hand columns on the matrix
// cores.  The augmented panel is the one of bcr_factor_generic.hip: R = 2n +
// nr rows [row jp; row j; border], columns [T = col j | col jp | col jpp | xb].
// Two launches per level, so that the latency-bound pivot steps run with
// many pairs per CU (A needs 2 KB of LDS) and only B holds the LDS a GEMM
// needs.
//
//  A. pair_lu_t: the T panel [E_jp; D_j; C_j] (R x n), one wave per pair,
//     in registers: rows over lanes (row lane + 64 s, s < 2), columns in
//     two (n <= 32) or three 16-wide register vectors; LAPACK partial pivoting among the first 2n
//     rows (first maximum), multiplier a * (1/pivot), fma updates -- the
//     arithmetic of the other LU kernels, so T, the pivots and C_j U^-1
//     agree with them bit for bit.  The step loop is rolled within each
//     16-column block (column c read and written through a uniform register
//     index; FP32: in groups of four steps, each updating its block's
//     columns from the group's first on).  Rows do not move; each row is
//     stored at its final position (T, C_j U^-1, perm), and the positions
//     themselves go to the unused tail of T (one byte per row).
//  B. pair_rhs_mfma: the right-hand part X = [col jp | col jpp | xb], two
//     waves per pair.  The row positions come from A's bytes; LDS gets
//     -L11 below its 16 x 16 diagonal blocks, the inverses of those unit
//     lower blocks on them (forward substitution, one column per lane), and
//     -[L21; C_j U^-1].  X goes in 16-column tiles, each inside one of the
//     three column blocks, its sources loaded one tile ahead and staged in
//     LDS with the rows at their final positions; then
//        U12_kb  = inv(L_kk) (X_kb - sum_{m<kb} L_km U12_m)
//                                           -> pivot rows: D_j', E_j', B_j'
//        X_rest' = X_rest - [L21; C_j U^-1] U12
//                                           -> reduced rows D_jp', E_jp', B_jp'
//     on v_mfma_f32_16x16x4f32 (FP32 factors) or v_mfma_f64_16x16x4f64 (FP64
//     factors).  The sum over k runs as 16-wide blocks, lane group
//     g = lane / 16 taking k = 16 kb + kro(g, i) in the i-th MFMA of a block
//     (kro: the row the result layout puts in value i of group g, 4 g + i in
//     FP32, g + 4 i in FP64): in FP32 every LDS operand is one 16-byte read
//     per lane per four MFMAs, and in both a block of U12 in the accumulator
//     layout is the next products' B operand as it stands.  The border rows
//     (C_jp', pending C_jpp, F partial) leave through LDS in one coalesced
//     pass per tile.
// The right-hand results differ from the rank-1 kernels in rounding only
// (other summation order, the inverted diagonal blocks).  A pivot whose
// reciprocal overflows (a subnormal pivot) turns the finished T columns into
// NaN here, where the rank-1 kernels leave them; the factorization is
// unusable either way and the refinement reports it.
#include "bcr_launch.hh"
#include "bcr_batch_kernels.hh"
#include "bcr_wave.hh"

namespace bcr {
namespace {

template <typename T>
using V2 = T __attribute__((ext_vector_type(2)));
template <typename T>
using V4 = T __attribute__((ext_vector_type(4)));
template <typename T>
using V16 = T __attribute__((ext_vector_type(16)));

// the most border rows (nr) the matrix-core pair takes
constexpr int kNrB = 8;

// LDS geometry: NT top slots (pivot positions, NMAX a multiple of 16), NR
// rest slots (reduced and border rows, n + nr <= RESTMAX)
template <int NMAX, int RESTMAX, int RS>
struct MfmaGeom {
  static constexpr int NT = NMAX;
  static constexpr int NR = (RESTMAX + 15) / 16 * 16;
  static constexpr int LD = NT + 4;        // row stride of Ls
  static constexpr int XLD = NT + NR + 4;  // column stride of a staged tile
  static_assert(NT % 16 == 0, "NMAX must be a multiple of 16");
  static_assert(NT + NR <= 64 * RS, "padding slots come from the lanes beyond row R");
};

// one 16 x 16 x 4 product-accumulate: FP32 or FP64 matrix cores (the same
// operand and accumulator layouts)
__device__ __forceinline__ V4<float> mfma1(float a, float b, V4<float> c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ V4<double> mfma1(double a, double b, V4<double> c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// The 16 x 16 x 4 products take A (16 x 4) as A[m][k] in lane m + 16 k and B
// (4 x 16) as B[k][n] in lane n + 16 k, in both precisions; the 16 x 16
// result comes back as four values per lane, lane n + 16 g holding in its
// value i row 4 g + i (FP32) or row g + 4 i (FP64) of column n.  kro gives
// that row; every k sum below runs in the same order (the i-th product of a
// 16-wide block takes k = kro(g, i) in lane group g), so a 16-row block in
// the result layout is the next product's B operand as it stands.
template <typename T>
__device__ __forceinline__ constexpr int kro(int g, int i) {
  return sizeof(T) == 4 ? 4 * g + i : g + 4 * i;
}
// rows kro(g, 0..3) of a 16-row block (LDS): one 16-byte read for FP32,
// four strided reads for FP64
template <typename T>
__device__ __forceinline__ V4<T> ld4(const T* blk, int g) {
  if constexpr (sizeof(T) == 4) {
    return *reinterpret_cast<const V4<T>*>(blk + 4 * g);
  } else {
    return V4<T>{blk[g], blk[g + 4], blk[g + 8], blk[g + 12]};
  }
}

template <typename T>
__device__ __forceinline__ V4<T> mfma4(V4<T> a, V4<T> b, V4<T> c) {
  c = mfma1(a.x, b.x, c);
  c = mfma1(a.y, b.y, c);
  c = mfma1(a.z, b.z, c);
  c = mfma1(a.w, b.w, c);
  return c;
}

// rows 16 kb + kro(g, i) (i < 4) of one column, those below lim: for FP32 one
// 16-byte store when the 16-row block lies below lim (a uniform test),
// single stores in the block lim cuts
template <typename T>
__device__ __forceinline__ void store_block(T* col, int kb, int g, V4<T> v, int lim) {
  if constexpr (sizeof(T) == 4) {
    const int r0 = 16 * kb + 4 * g;
    if (16 * kb + 16 <= lim) {
      __builtin_memcpy(col + r0, &v, 4 * sizeof(T));   // element-aligned; the hardware takes it
    } else if (16 * kb < lim) {
#pragma unroll
      for (int i = 0; i < 4; ++i)
        if (r0 + i < lim) col[r0 + i] = v[i];
    }
  } else {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int r = 16 * kb + kro<T>(g, i);
      if (r < lim) col[r] = v[i];
    }
  }
}

// v when the mask is all ones, +0 when it is zero
__device__ __forceinline__ float mask_bits(float v, int m) { return __int_as_float(__float_as_int(v) & m); }
__device__ __forceinline__ double mask_bits(double v, int m) {
  return __longlong_as_double(__double_as_longlong(v) & static_cast<long long>(m));
}

// lane `who`'s value of v when K > cc, else 0: a scalar select on the
// v_readlane result (written out: the compiler would move the value to a
// VGPR and select there, two vector instructions instead of none)
template <int K>
__device__ __forceinline__ float rl_after(float v, int who, int cc) {
  int x = __builtin_amdgcn_readlane(__float_as_int(v), who);
  asm("s_cmp_lt_u32 %1, %2\n\ts_cselect_b32 %0, %0, 0" : "+s"(x) : "s"(cc), "n"(K) : "scc");
  return __int_as_float(x);
}
template <int K>
__device__ __forceinline__ double rl_after(double v, int who, int cc) {
  const long long b = __double_as_longlong(v);
  int lo = __builtin_amdgcn_readlane(int(b & 0xffffffff), who);
  int hi = __builtin_amdgcn_readlane(int(b >> 32), who);
  asm("s_cmp_lt_u32 %2, %3\n\ts_cselect_b32 %0, %0, 0\n\ts_cselect_b32 %1, %1, 0"
      : "+s"(lo), "+s"(hi) : "s"(cc), "n"(K) : "scc");
  return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}

// columns K.. 15 of one 16-column block of the row sets, those not after
// column cc of the block getting a zero pivot-row entry; the pivot row is in
// row set 0 or 1 (a third row set holds border rows only)
// f(integral_constant<int, I>) for I = B .. E-1, unrolled at compile time
template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

// Columns before KM may lie at or before cc and are masked; the later ones
// always lie after it.
template <int K, int KM, int RS, typename TW>
__device__ __forceinline__ void update_after(V16<TW>& v0, V16<TW>& v1, V16<TW>& v2, bool sp, int who, int cc, TW lv0,
                                             TW lv1, TW lv2) {
  if constexpr (K < 16) {
    const TW u = K < KM ? rl_after<K>(sp ? v1[K] : v0[K], who, cc) : wave::rl(sp ? v1[K] : v0[K], who);
    v0[K] = fma(-lv0, u, v0[K]);
    v1[K] = fma(-lv1, u, v1[K]);
    if constexpr (RS > 2) v2[K] = fma(-lv2, u, v2[K]);
    update_after<K + 1, KM, RS, TW>(v0, v1, v2, sp, who, cc, lv0, lv1, lv2);
  }
}

// X_rest' = X_rest - [L21; C_j U^-1] U12 over the first NRT 16-row tiles of
// the rest rows (the minus sign is in Ls); the tiles' independent
// accumulators alternate on the matrix core.  Reduced rows go to drest,
// border rows back to their staging slots.
template <int NB, int NRT, int NT, int LD, int XLD, typename TW>
__device__ __forceinline__ void rest_tiles(const V4<TW> (&u)[NB], const TW* Ls, TW* xt, int cl, int g, int n, int nr,
                                           bool colok, TW* drest) {
  V4<TW> acc[NRT];
#pragma unroll
  for (int rt = 0; rt < NRT; ++rt) acc[rt] = ld4(xt + cl * XLD + NT + rt * 16, g);
#pragma unroll
  for (int kb = 0; kb < NB; ++kb) {
    V4<TW> av[NRT];
#pragma unroll
    for (int rt = 0; rt < NRT; ++rt)
      av[rt] = ld4(Ls + (NT + rt * 16 + cl) * LD + kb * 16, g);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int rt = 0; rt < NRT; ++rt) acc[rt] = mfma1(av[rt][i], u[kb][i], acc[rt]);
  }
  // rest position q = 16 rt + kro(g, i): reduced row q (< n), else border row q - n
  if (colok) {
#pragma unroll
    for (int rt = 0; rt < NRT; ++rt) store_block(drest, rt, g, acc[rt], n);
  }
#pragma unroll
  for (int rt = 0; rt < NRT; ++rt)
    if (16 * rt + 16 > n && 16 * rt < n + nr) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int q = rt * 16 + kro<TW>(g, i);
        if (q >= n && q < n + nr) xt[cl * XLD + NT + q] = acc[rt][i];
      }
    }
}

// A panel row of the right-hand part: its class (1: a row of block jp, 2: a
// row of block j, 4: a border row, 0: none) and its row in that block.
struct RowSrc {
  int cls, off;
};

// the raw source values of one row in right-hand tile t (16 columns of the
// tile's region; clamped addresses, masked when staged)
// (the source blocks come as values read from the argument record before
// the kernel's first store: a batched kernel's record sits in memory, and a
// field read after a store is fetched again, behind it)
template <typename TS>
__device__ __forceinline__ void load_rhs_row(TS (&xs)[16], const TS* sD, const TS* sE, const TS* sB, const TS* sC,
                                             int n, int nr, int nx, int t, int tA, int j, int jp, RowSrc rs) {
  const size_t nn = size_t(n) * n, rn = size_t(nr) * n, bn = size_t(n) * nx;
  const int reg = t < tA ? 0 : t < 2 * tA ? 1 : 2;
  const int c0 = 16 * (reg == 0 ? t : reg == 1 ? t - tA : t - 2 * tA);
  const int width = reg == 2 ? nx : n;
  // region 0: [D_jp | . | C_jp]; region 1: [. | E_j | .]; region 2: [B_jp | B_j | .]
  const bool valid = reg == 0 ? (rs.cls & 5) != 0 : reg == 1 ? rs.cls == 2 : (rs.cls & 3) != 0;
  const TS* src = reg == 0 ? (rs.cls == 4 ? sC + jp * rn : sD + jp * nn)
                : reg == 1 ? sE + j * nn
                           : sB + (rs.cls == 1 ? jp : j) * bn;
  src = valid ? src + rs.off : sD + jp * nn;   // a valid address for the rows of zeros
  const int stride = reg == 0 && rs.cls == 4 ? nr : n;
#pragma unroll
  for (int cc = 0; cc < 16; ++cc) xs[cc] = src[(c0 + cc < width ? c0 + cc : width - 1) * stride];
}

// ---- A: the T panel, one wave per pair ------------------------------------------
template <typename TS, typename TW, int NB, int RS, int NRMAX>
__device__ __forceinline__ void pair_lu_t_body(FactorArgs<TS, TW> const& a, int p) {
  constexpr int NMAX = 16 * NB;
  static_assert(RS == 2 || RS == 3, "two or three row sets");
  static_assert(NMAX <= 64, "the 2n candidate rows lie in row sets 0 and 1");
  __shared__ __align__(16) TW stage[(NRMAX + 1) * NMAX];   // border rows C_j, one zero row
  if (p >= a.npairs) return;
  const int lane = threadIdx.x & 63;
  const int n = a.n, nr = a.nr, n2 = 2 * n;
  const int R = n2 + nr;
  int j, jp, jpp;
  pair_of_uniform(a, p, j, jp, jpp);
  const size_t nn = size_t(n) * n, rn = size_t(nr) * n;
  const TW* pend_cur = a.Cpend + size_t(a.slot_cur) * (a.N + 1) * rn;
  const TW* pend_prv = a.Cpend + size_t(a.slot_cur ^ 1) * (a.N + 1) * rn;

  // the border rows C_j (+ pending C_j) first, coalesced, into LDS rows of
  // NMAX (row nr stays zero for the other lanes)
  for (int e = lane; e < (nr + 1) * NMAX; e += 64) stage[e] = TW(0);
  for (int e = lane; e < nr * n; e += 64) {   // C_j is nr x n, column-major
    const int col = e / nr, ri = e - col * nr;
    TW x = TW(a.sC[j * rn + e]);
    if (!a.first) x += pend_prv[j * rn + e] + pend_cur[j * rn + e];
    stage[ri * NMAX + col] = x;
  }
  V16<TW> V[RS][NB];
#pragma unroll
  for (int s = 0; s < RS; ++s) {
    const int r = lane + 64 * s;
    const bool top = r < n, mid = r >= n && r < n2, bord = r >= n2 && r < R;
    // rows of block jp: E_jp; rows of block j: D_j
    const TS* base = top ? a.sE + jp * nn + r : a.sD + j * nn + (mid ? r - n : 0);
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        const int col = 16 * b + k, off = (col < n ? col : n - 1) * n;
        V[s][b][k] = keep_if(TW(base[off]), (top || mid) && col < n);
      }
    if (R > 64 * s && n2 < 64 * (s + 1)) {   // this row set holds border rows
      const TW* srow = stage + (bord ? r - n2 : nr) * NMAX;
#pragma unroll
      for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int k = 0; k < 16; k += 4) {
          const V4<TW> c4 = *reinterpret_cast<const V4<TW>*>(srow + 16 * b + k);
          V[s][b][k] += c4.x;
          V[s][b][k + 1] += c4.y;
          V[s][b][k + 2] += c4.z;
          V[s][b][k + 3] += c4.w;
        }
    }
  }
  int pos0 = lane, pos1 = lane + 64;   // positions of rows lane, lane + 64
  // 2n <= 64: every pivot candidate sits in row set 0 (row set 1 holds
  // border rows only), and the pivot row's entries come from it unselected
  constexpr bool ONE = 2 * NMAX <= 64;
  int myperm = 0;
  bool zero_pivot = false;
  // Steps in blocks of 16, each in GR groups of GS steps: step c of block b
  // reads and rewrites column c through a uniform register index
  // (s_set_gpr_idx) and updates the columns of blocks b.. with the pivot
  // row's entries; group g of block b updates its block's columns from
  // GS g + 1 on, those that may not lie after c getting a zero entry.  Each
  // group's loop stays rolled (one copy of a step per group).  FP32: four
  // groups (a quarter of the masked in-block updates); FP64: one (the
  // copies would cost the FP64 panel a wave per SIMD).
  constexpr int GR = sizeof(TW) == 4 ? 4 : 1, GS = 16 / GR;
  // step c of block b; the block's columns from KS on are updated, those
  // before KM masked
  auto step = [&](const int b, const int c, auto KS, auto KM) {
    const int cc = c - 16 * b;
    const TW x0 = V[0][b][cc], x1 = V[1][b][cc];
    const TW x2 = RS > 2 ? V[RS - 1][b][cc] : TW(0);   // border rows (row set 2)
    const bool c0 = lane < n2 && pos0 >= c;
    const bool c1 = !ONE && lane + 64 < n2 && pos1 >= c;
    unsigned h0, l0, h1, l1;
    wave::abs_key(x0, h0, l0);
    wave::abs_key(x1, h1, l1);
    // the better of this lane's two candidates (ties to the smaller position)
    const bool b1 = c1 & (!c0 | (h1 > h0) | ((h1 == h0) & ((l1 > l0) | ((l1 == l0) & (pos1 < pos0)))));
    bool zero;
    const int who = wave::argmax_pos<(sizeof(TW) > 4)>(c0 | c1, b1 ? h1 : h0, b1 ? l1 : l0, b1 ? pos1 : pos0, lane,
                                                        zero);
    const int P = __builtin_amdgcn_readlane(b1 ? pos1 : pos0, who);
    const bool sp = !ONE && __builtin_amdgcn_readlane(int(b1), who) != 0;
    zero_pivot = zero_pivot || zero;
    myperm = lane == c ? P + 1 : myperm;
    // the pivot row moves to position c, the row at c to the pivot's old place
    pos0 = pos0 == c ? P : pos0 == P ? c : pos0;
    if constexpr (!ONE) pos1 = pos1 == c ? P : pos1 == P ? c : pos1;
    const TW pv = wave::rl(sp ? x1 : x0, who);
    const TW rinv = pv != TW(0) ? TW(1) / pv : TW(0);
    // rows below the pivot get the multiplier, the others keep column c
    // (border and padding rows hold positions r >= 2n > c: always below;
    // the padding rows are zero and never stored)
    const bool u0 = pos0 > c;
    const bool u1 = pos1 > c;
    const TW m0 = x0 * rinv, m1 = x1 * rinv;
    const TW lv0 = u0 ? m0 : TW(0), lv1 = u1 ? m1 : TW(0);
    V[0][b][cc] = u0 ? m0 : x0;
    V[1][b][cc] = u1 ? m1 : x1;
    TW lv2 = TW(0);
    if constexpr (RS > 2) {   // border (and zero padding) rows: never pivots, always updated
      lv2 = x2 * rinv;
      V[RS - 1][b][cc] = lv2;
    }
    // columns KS .. 15 of the block (those before are never after c)
    update_after<decltype(KS)::value, decltype(KM)::value, RS, TW>(V[0][b], V[1][b], V[RS - 1][b], sp, who, cc, lv0,
                                                                   lv1, lv2);
#pragma unroll
    for (int bb = b + 1; bb < NB; ++bb)
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        const TW u = wave::rl(sp ? V[1][bb][k] : V[0][bb][k], who);
        V[0][bb][k] = fma(-lv0, u, V[0][bb][k]);
        V[1][bb][k] = fma(-lv1, u, V[1][bb][k]);
        if constexpr (RS > 2) V[RS - 1][bb][k] = fma(-lv2, u, V[RS - 1][bb][k]);
      }
  };
  if constexpr (GR == 1) {
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int cend = min(16 * b + 16, n);
      for (int c = 16 * b; c < cend; ++c) step(b, c, std::integral_constant<int, 1>{}, std::integral_constant<int, 16>{});
    }
  } else {
    static_for<0, GR * NB>([&](auto BG) {
      constexpr int b = decltype(BG)::value / GR, g = decltype(BG)::value % GR;
      const int cend = min(16 * b + GS * g + GS, n);
      for (int c = 16 * b + GS * g; c < cend; ++c)
        step(b, c, std::integral_constant<int, GS * g + 1>{}, std::integral_constant<int, GS * g + GS>{});
    });
  }
  if (zero_pivot && lane == 0) atomicOr(a.status, 1);
  // T (2n x n, final row order) and C_j U^-1 to their blocks
  TW* Tj = a.Tm + size_t(j) * tsize(n);
#pragma unroll
  for (int s = 0; s < RS; ++s) {
    const int r = lane + 64 * s;
    const int q = r < n2 ? (s == 0 ? pos0 : pos1) : r;
    if (r < R) {
      TW* dst = q < n2 ? Tj + q : a.C + j * rn + (q - n2);
      const int stride = q < n2 ? n2 : nr;
#pragma unroll
      for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int k = 0; k < 16; ++k)
          if (16 * b + k < n) dst[size_t(16 * b + k) * stride] = V[s][b][k];
    }
  }
  if (lane < n) a.perm[size_t(j) * n + lane] = myperm;
  // the final position of every panel row, one byte each after the panel's
  // 2n^2 entries (the LU pair leaves that tail of T unused): kernel B reads
  // them instead of replaying the swap sequence
  uint8_t* const posb = reinterpret_cast<uint8_t*>(Tj + size_t(2) * n * n);
  if (lane < n2) posb[lane] = uint8_t(pos0);
  if (lane + 64 < n2) posb[lane + 64] = uint8_t(pos1);
}

// ---- B: the right-hand part on the matrix cores ----------------------------------
template <typename TS, typename TW, int NB, int RS, int RESTMAX, int NW>
__device__ __forceinline__ void pair_rhs_body(FactorArgs<TS, TW> const& a, int p) {
  using G = MfmaGeom<16 * NB, RESTMAX, RS>;
  constexpr int NT = G::NT, NR = G::NR, LD = G::LD, XLD = G::XLD;
  __shared__ __align__(16) TW Ls[(NT + NR) * LD];
  __shared__ __align__(16) TW Xt[NW][16 * XLD];
  __shared__ unsigned sstat[NW][3];   // per-wave panel statistics
  if (p >= a.npairs) return;   // uniform over the workgroup
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
  const int n = a.n, nr = a.nr, nx = a.nx, n2 = 2 * n;
  const int R = n2 + nr;
  int j, jp, jpp;
  pair_of_uniform(a, p, j, jp, jpp);
  const size_t nn = size_t(n) * n, rn = size_t(nr) * n, bn = size_t(n) * nx;
  TW* pend_cur = a.Cpend + size_t(a.slot_cur) * (a.N + 1) * rn;
  TW* pend_prv = a.Cpend + size_t(a.slot_cur ^ 1) * (a.N + 1) * rn;
  const TW* Tj = a.Tm + size_t(j) * tsize(n);
  // the record's pointers, read once before the first store (a batched
  // kernel's record sits in memory: a field read after a store is fetched
  // again, a dependent memory round trip in front of the load it addresses)
  const TS *const sD = a.sD, *const sE = a.sE, *const sB = a.sB, *const sC = a.sC;
  TW *const Dm = a.D, *const Em = a.E, *const Bm = a.B, *const Cm = a.C, *const Fm = a.Fpart;
  // the final positions of this lane's two panel rows, stored by kernel A
  // after the panel's 2n^2 entries
  const uint8_t* const posb = reinterpret_cast<const uint8_t*>(Tj + size_t(2) * n * n);
  const uint8_t pb0 = posb[lane < n2 ? lane : 0], pb1 = posb[lane + 64 < n2 ? lane + 64 : 0];

  // this lane's two panel rows as right-hand sources, and the first tile's
  // loads, in flight while the multipliers are staged
  RowSrc rsrc[RS];
#pragma unroll
  for (int s = 0; s < RS; ++s) {
    const int r = lane + 64 * s;
    rsrc[s].cls = r < n ? 1 : r < n2 ? 2 : r < R ? 4 : 0;
    rsrc[s].off = r < n ? r : r < n2 ? r - n : r < R ? r - n2 : 0;
  }
  const int tA = (n + 15) / 16, tC = (nx + 15) / 16;
  const int ntiles = 2 * tA + tC;
  // the tile's raw source values, loaded one tile ahead: row lane + 64 s,
  // column c0 + cc of the tile's region
  TS xs[RS][16];
  if (w < ntiles) {
#pragma unroll
    for (int s = 0; s < RS; ++s) load_rhs_row(xs[s], sD, sE, sB, sC, n, nr, nx, w, tA, j, jp, rsrc[s]);
  }

  // Ls: top rows (final positions < n) hold -L11 below the diagonal blocks
  // and the inverses of L11's unit lower 16 x 16 diagonal blocks on them
  // (set below); rest rows hold -[L21; C_j U^-1].  Rows are final positions,
  // so every read is a column of T (or of C_j U^-1).
  // Lanes over the columns k of T (column-major: a lane reads along its
  // column, and the writes of Ls rows are bank-conflict-free); all loads are
  // clamped and issued before their first use (T blocks are 4-byte aligned
  // only: n (2n + 1) floats each).  Top row q < n: L below the diagonal, negated
  // outside the diagonal blocks; rest row n + i: -T row n + i; border rows
  // and padding rows below.
  {
    const int k = lane, kk = k < n ? k : n - 1;
    constexpr int PAIRS = (2 * NT + 2 * NW - 1) / (2 * NW);   // row pairs of T per wave
    V2<TW> tv[PAIRS];
#pragma unroll
    for (int h = 0; h < PAIRS; ++h) {
      const int q0 = 2 * (w + NW * h), qc = q0 < n2 ? q0 : 0;
      tv[h].x = Tj[qc + size_t(kk) * n2];
      tv[h].y = Tj[qc + 1 + size_t(kk) * n2];
    }
    // the border rows C_j U^-1 (nr <= kNrB, nr x n column-major), issued with
    // the T loads; wave 0 stores them
    TW cv[kNrB];
    if (nr > 0) {
#pragma unroll
      for (int ri = 0; ri < kNrB; ++ri) cv[ri] = Cm[j * rn + (ri < nr ? ri : 0) + size_t(kk) * nr];
    }
    // the selector's statistics of this panel (FP32 batches): |U_qq| min and
    // max, |U_qk| max over q <= k, as float bits (non-negative: ordered as
    // unsigned), from the U entries this wave reads anyway
    const bool stats = sizeof(TW) == 4 && a.pstat != nullptr;
    unsigned dmin = 0xffffffffu, dmax = 0u, umax = 0u;
#pragma unroll
    for (int h = 0; h < PAIRS; ++h) {
      const int q0 = 2 * (w + NW * h);
      if (q0 < n2 && k < NT) {
#pragma unroll
        for (int d = 0; d < 2; ++d) {
          const int q = q0 + d;
          const TW v = d == 0 ? tv[h].x : tv[h].y;
          TW x;
          int row;
          if (q < n) {
            row = q;
            x = k < q ? ((k >> 4) == (q >> 4) ? v : -v) : TW(0);
            if (stats && k >= q && k < n) {
              const unsigned bits = __float_as_uint(fabsf(float(v)));
              umax = max(umax, bits);
              if (k == q) {
                dmin = min(dmin, bits);
                dmax = max(dmax, bits);
              }
            }
          } else {
            row = NT + q - n;
            x = -v;
          }
          Ls[row * LD + k] = keep_if(x, k < n);
        }
      }
    }
    if (stats) {
      dmin = wave::min_u32(dmin);
      dmax = wave::max_u32(dmax);
      umax = wave::max_u32(umax);
      if (lane == 0) {
        sstat[w][0] = dmin;
        sstat[w][1] = dmax;
        sstat[w][2] = umax;
      }
    }
    // border rows (C_j U^-1, nr x n column-major) and the zero padding rows
    if (w == 0 && k < NT) {
#pragma unroll
      for (int ri = 0; ri < kNrB; ++ri)
        if (ri < nr) Ls[(NT + n + ri) * LD + k] = keep_if(-cv[ri], k < n);
      for (int q = n; q < NT; ++q) Ls[q * LD + k] = TW(0);
      for (int i = n + nr; i < NR; ++i) Ls[(NT + i) * LD + k] = TW(0);
    }
  }
  // the final position of this lane's two rows (kernel A's bytes)
  const int pos0 = lane < n2 ? int(pb0) : lane, pos1 = lane + 64 < n2 ? int(pb1) : lane + 64;
  __syncthreads();   // Ls loaded
  if (sizeof(TW) == 4 && a.pstat != nullptr && w == 0 && lane == 0) {
    unsigned dmin = sstat[0][0], dmax = sstat[0][1], umax = sstat[0][2];
#pragma unroll
    for (int v = 1; v < NW; ++v) {
      dmin = min(dmin, sstat[v][0]);
      dmax = max(dmax, sstat[v][1]);
      umax = max(umax, sstat[v][2]);
    }
    float* const o = a.pstat + 3 * size_t(j - 1);
    o[0] = __uint_as_float(dmin);
    o[1] = __uint_as_float(dmax);
    o[2] = __uint_as_float(umax);
  }
  if (lane < NT && (lane >> 4) % NW == w) {
    // the inverse of diagonal block lane / 16 of L11 (unit lower), column
    // lane % 16, by forward substitution in place (the blocks dealt out to
    // the waves); rows beyond n are zero, their inverse rows identity
    const int b = lane >> 4, jj = lane & 15;
    TW y[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) y[i] = TW(0);
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const TW* li = Ls + (16 * b + i) * LD + 16 * b;
      TW t0 = TW(0), t1 = TW(0);
#pragma unroll
      for (int k0 = 0; k0 < i; k0 += 4) {
        const V4<TW> l4 = *reinterpret_cast<const V4<TW>*>(li + k0);
        t0 = fma(l4.x, y[k0], t0);
        if (k0 + 1 < 16) t1 = fma(l4.y, y[k0 + 1 < 16 ? k0 + 1 : 0], t1);
        if (k0 + 2 < 16) t0 = fma(l4.z, y[k0 + 2 < 16 ? k0 + 2 : 0], t0);
        if (k0 + 3 < 16) t1 = fma(l4.w, y[k0 + 3 < 16 ? k0 + 3 : 0], t1);
      }
      y[i] = jj == i ? TW(1) : -(t0 + t1);
    }
    // every lane has read its block rows (reads precede these writes in this wave)
#pragma unroll
    for (int i = 0; i < 16; ++i) Ls[(16 * b + i) * LD + 16 * b + jj] = y[i];
  }
  __syncthreads();   // the diagonal inverses in place

  TW* xt = Xt[w];
  const int g = lane >> 4, cl = lane & 15;
  int slot[RS];
#pragma unroll
  for (int s = 0; s < RS; ++s) {
    const int r = lane + 64 * s;
    if (r < n2) {
      const int q = s == 0 ? pos0 : pos1;
      slot[s] = q < n ? q : NT + q - n;
    } else if (r < R) {
      slot[s] = NT + r - n;
    } else {   // lanes beyond the panel zero the padding slots
      const int i = r - R;
      slot[s] = i < NT - n ? n + i : i - (NT - n) < NR - (n + nr) ? NT + n + nr + (i - (NT - n)) : NT + NR;
    }
  }
  for (int t = w; t < ntiles; t += NW) {
    const int reg = t < tA ? 0 : t < 2 * tA ? 1 : 2;
    const int c0 = 16 * (reg == 0 ? t : reg == 1 ? t - tA : t - 2 * tA);
    const int width = reg == 2 ? nx : n;
    // stage the tile: column cc at xt[cc * XLD + slot], zeros outside the region's rows
#pragma unroll
    for (int s = 0; s < RS; ++s) {
      const int cls = rsrc[s].cls;
      const bool valid = reg == 0 ? (cls & 5) != 0 : reg == 1 ? cls == 2 : (cls & 3) != 0;
      // a whole tile (a uniform test): one row mask for all columns (not for
      // NB = 2, where the second copy of the loop would cost a wave per SIMD)
      // (FP32 only: the integer mask is one 32-bit AND per value)
      if (sizeof(TW) == 4 && NB > 2 && c0 + 16 <= width) {
        int m = valid ? -1 : 0;
        asm volatile("" : "+v"(m));
#pragma unroll
        for (int cc = 0; cc < 16; ++cc) xt[cc * XLD + slot[s]] = mask_bits(TW(xs[s][cc]), m);
      } else {
#pragma unroll
        for (int cc = 0; cc < 16; ++cc) xt[cc * XLD + slot[s]] = keep_if(TW(xs[s][cc]), valid && c0 + cc < width);
      }
    }
    // the border rows' extra terms, nr x 16 contiguous values each (element
    // e = border row e % nr, tile column e / nr): pending C_jp (region 0, read
    // once, cleared for reuse) and the pending C_jpp this tile adds to (region 1)
    const int m = nr * min(16, width - c0);
    const size_t eb = size_t(c0) * nr;
    TW pin[2] = {TW(0), TW(0)};
    if ((reg == 0 && !a.first) || reg == 1) {
      const TW* pe = reg == 0 ? pend_prv + jp * rn + eb : pend_cur + jpp * rn + eb;
#pragma unroll
      for (int h = 0; h < 2; ++h)
        if (lane + 64 * h < m) pin[h] = pe[lane + 64 * h];
    }
    // the next tile's loads fly over this tile's work
    if (t + NW < ntiles) {
#pragma unroll
      for (int s = 0; s < RS; ++s) load_rhs_row(xs[s], sD, sE, sB, sC, n, nr, nx, t + NW, tA, j, jp, rsrc[s]);
    }
    if (reg == 0 && !a.first) {   // pending C_jp into the border slots
      TW* pe = pend_prv + jp * rn + eb;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int e = lane + 64 * h;
        if (e < m) {
          const int cc = e / nr, rb = e - cc * nr;
          xt[cc * XLD + NT + n + rb] += pin[h];
          pe[e] = TW(0);
        }
      }
    }

    // where this lane's column goes: pivot rows -> block j, reduced rows -> block jp
    const int kc = c0 + cl;
    const bool colok = kc < width;
    TW* const blk = reg == 0 ? Dm : reg == 1 ? Em : Bm;
    const size_t bsz = reg == 2 ? bn : nn;
    const int kcs = colok ? kc : 0;
    TW* const dtop = blk + size_t(j) * bsz + size_t(kcs) * n;
    TW* const drest = blk + size_t(jp) * bsz + size_t(kcs) * n;

    // U12 = L11^-1 X_top by 16-row blocks: u_kb = inv(L_kk) (X_kb - sum_{m<kb}
    // L_km u_m).  A block of U12 in the accumulator layout (rows 4g + i,
    // column cl) is the B operand layout of the next products as it stands.
    V4<TW> u[NB];
#pragma unroll
    for (int kb = 0; kb < NB; ++kb) {
      V4<TW> acc = ld4(xt + cl * XLD + kb * 16, g);
#pragma unroll
      for (int mb = 0; mb < kb; ++mb)   // -L_km is in Ls
        acc = mfma4<TW>(ld4(Ls + (kb * 16 + cl) * LD + mb * 16, g), u[mb], acc);
      u[kb] = mfma4<TW>(ld4(Ls + (kb * 16 + cl) * LD + kb * 16, g), acc, V4<TW>{TW(0), TW(0), TW(0), TW(0)});
    }
    if (colok) {
#pragma unroll
      for (int kb = 0; kb < NB; ++kb) store_block(dtop, kb, g, u[kb], n);
    }
    // X_rest' = X_rest - [L21; C_j U^-1] U12 (the minus sign is in Ls), over
    // the 16-row tiles the n + nr rest rows occupy (one fewer than NR / 16
    // when n + nr <= NR - 16)
    if (n + nr <= NR - 16)
      rest_tiles<NB, NR / 16 - 1, NT, LD, XLD, TW>(u, Ls, xt, cl, g, n, nr, colok, drest);
    else
      rest_tiles<NB, NR / 16, NT, LD, XLD, TW>(u, Ls, xt, cl, g, n, nr, colok, drest);
    // border rows: C_jp' (region 0), pending C_jpp += (region 1), F partial (region 2)
    TW* const bdst = reg == 0 ? Cm + jp * rn + eb : reg == 1 ? pend_cur + jpp * rn + eb
                                                              : Fm + size_t(j) * nr * nx + eb;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int e = lane + 64 * h;
      if (e < m) {
        const int cc = e / nr, rb = e - cc * nr;
        const TW v = xt[cc * XLD + NT + n + rb];
        bdst[e] = reg == 1 ? pin[h] + v : v;
      }
    }
  }
}

template <typename TS, typename TW, int NB, int RS, int NRMAX>
__global__ void __launch_bounds__(64) pair_lu_t_kernel(FactorArgs<TS, TW> a) {
  pair_lu_t_body<TS, TW, NB, RS, NRMAX>(a, blockIdx.x);
}
template <typename TS, typename TW, int NB, int RS, int NRMAX>
__global__ void __launch_bounds__(64) pair_lu_t_batched(BatchTable<FactorArgs<TS, TW>> t) {
  const int e = t.entry[blockIdx.x];
  pair_lu_t_body<TS, TW, NB, RS, NRMAX>(t.args[e], int(blockIdx.x) - t.first[e]);
}
template <typename TS, typename TW, int NB, int RS, int RESTMAX, int NW>
__global__ void __launch_bounds__(64 * NW) pair_rhs_mfma_kernel(FactorArgs<TS, TW> a) {
  pair_rhs_body<TS, TW, NB, RS, RESTMAX, NW>(a, blockIdx.x);
}
template <typename TS, typename TW, int NB, int RS, int RESTMAX, int NW>
__global__ void __launch_bounds__(64 * NW) pair_rhs_mfma_batched(BatchTable<FactorArgs<TS, TW>> t) {
  const int e = t.entry[blockIdx.x];
  pair_rhs_body<TS, TW, NB, RS, RESTMAX, NW>(t.args[e], int(blockIdx.x) - t.first[e]);
}

// waves per pair in the right-hand kernel
constexpr int kRhsWaves = 2;

// shapes: 1: n <= 32, n + nr <= 48 (two 16-column blocks); 2: n <= 48,
// n + nr <= 64 (three); 3: n <= 64, n + nr <= 80 (four, and a third row set
// for the border rows beyond row 128); all nr <= 8, nx <= 16; 0: not covered
int mfma_variant(int n, int nr, int nx) {
  if (n < 1 || nr > kNrB || nx > 16) return 0;
  if (n <= 32 && n + nr <= 48) return 1;
  if (n <= 48 && n + nr <= 64) return 2;
  if (n <= 64 && n + nr <= 80) return 3;
  return 0;
}

}  // namespace

bool pair_lu_mfma_fits(int n, int nr, int nx, bool fp64) {
  const int v = mfma_variant(n, nr, nx);
  return fp64 ? v == 1 || v == 2 : v != 0;
}

namespace {
// the A and B launches of shape variant V (FP64 factors: V = 1, 2; the T
// panel of V = 3 would need 384 registers per lane in FP64)
template <int V, typename TS, typename TW>
void launch_pair(FactorArgs<TS, TW> const& a, hipStream_t st) {
  constexpr int NB = V + 1, RS = V == 3 ? 3 : 2, RESTMAX = 32 + 16 * V;
  hipLaunchKernelGGL((pair_lu_t_kernel<TS, TW, NB, RS, kNrB>), dim3(a.npairs), dim3(64), 0, st, a);
  hipLaunchKernelGGL((pair_rhs_mfma_kernel<TS, TW, NB, RS, RESTMAX, kRhsWaves>), dim3(a.npairs), dim3(64 * kRhsWaves),
                     0, st, a);
}
template <int V, typename TS, typename TW>
void launch_pair_batched(BatchTable<FactorArgs<TS, TW>> const& t, int nblocks, hipStream_t st) {
  constexpr int NB = V + 1, RS = V == 3 ? 3 : 2, RESTMAX = 32 + 16 * V;
  hipLaunchKernelGGL((pair_lu_t_batched<TS, TW, NB, RS, kNrB>), dim3(nblocks), dim3(64), 0, st, t);
  hipLaunchKernelGGL((pair_rhs_mfma_batched<TS, TW, NB, RS, RESTMAX, kRhsWaves>), dim3(nblocks), dim3(64 * kRhsWaves),
                     0, st, t);
}
}  // namespace

template <typename TS, typename TW>
bool launch_pair_lu_mfma(FactorArgs<TS, TW> const& a, hipStream_t st, hipError_t* err) {
  constexpr bool f64 = sizeof(TW) == 8;
  switch (mfma_variant(a.n, a.nr, a.nx)) {
    case 1: launch_pair<1>(a, st); break;
    case 2: launch_pair<2>(a, st); break;
    case 3:
      if constexpr (f64) return false;
      else launch_pair<3>(a, st);
      break;
    default:
      return false;
  }
  *err = hipGetLastError();
  return true;
}

template <typename TS, typename TW>
bool launch_pair_lu_mfma_batched(BatchTable<FactorArgs<TS, TW>> const& t, int nblocks, int n, int nx, int nr,
                                 hipStream_t st, hipError_t* err) {
  constexpr bool f64 = sizeof(TW) == 8;
  switch (mfma_variant(n, nr, nx)) {
    case 1: launch_pair_batched<1>(t, nblocks, st); break;
    case 2: launch_pair_batched<2>(t, nblocks, st); break;
    case 3:
      if constexpr (f64) return false;
      else launch_pair_batched<3>(t, nblocks, st);
      break;
    default:
      return false;
  }
  *err = hipGetLastError();
  return true;
}

template bool launch_pair_lu_mfma<float, float>(FactorArgs<float, float> const&, hipStream_t, hipError_t*);
template bool launch_pair_lu_mfma<double, float>(FactorArgs<double, float> const&, hipStream_t, hipError_t*);
template bool launch_pair_lu_mfma<double, double>(FactorArgs<double, double> const&, hipStream_t, hipError_t*);
template bool launch_pair_lu_mfma_batched<float, float>(BatchTable<FactorArgs<float, float>> const&, int, int, int, int,
                                                        hipStream_t, hipError_t*);
template bool launch_pair_lu_mfma_batched<double, float>(BatchTable<FactorArgs<double, float>> const&, int, int, int,
                                                         int, hipStream_t, hipError_t*);
template bool launch_pair_lu_mfma_batched<double, double>(BatchTable<FactorArgs<double, double>> const&, int, int,
                                                          int, int, hipStream_t, hipError_t*);

}  // namespace bcr
