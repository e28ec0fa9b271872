// Wavefront-level helpers shared by the register-resident pair kernels:
// DPP max / min reductions over the 64 lanes, v_readlane broadcasts of one
// lane's value, and the integer key of |x| used by the pivot searches.
#pragma once
#include <hip/hip_runtime.h>

namespace bcr {
namespace wave {

// gfx9 DPP controls
constexpr int kQuad1032 = 0xB1, kQuad2301 = 0x4E, kHalfMirror = 0x141, kMirror = 0x140;
constexpr int kBcast15 = 0x142, kBcast31 = 0x143;

// x from the DPP source lane; lanes the row mask leaves out get `id` (the
// identity of the reduction), so the move can fold into the max / min
template <int CTRL, int ROWMASK = 0xf>
__device__ __forceinline__ unsigned mov(unsigned x, unsigned id) {
  return unsigned(__builtin_amdgcn_update_dpp(int(id), int(x), CTRL, ROWMASK, 0xf, false));
}

// max over the 64 lanes, returned uniform
__device__ __forceinline__ unsigned max_u32(unsigned x) {
  x = max(x, mov<kQuad1032>(x, 0u));
  x = max(x, mov<kQuad2301>(x, 0u));
  x = max(x, mov<kHalfMirror>(x, 0u));
  x = max(x, mov<kMirror>(x, 0u));
  x = max(x, mov<kBcast15, 0xa>(x, 0u));
  x = max(x, mov<kBcast31, 0xc>(x, 0u));
  return unsigned(__builtin_amdgcn_readlane(int(x), 63));
}

__device__ __forceinline__ unsigned min_u32(unsigned x) {
  constexpr unsigned id = 0xffffffffu;
  x = min(x, mov<kQuad1032>(x, id));
  x = min(x, mov<kQuad2301>(x, id));
  x = min(x, mov<kHalfMirror>(x, id));
  x = min(x, mov<kMirror>(x, id));
  x = min(x, mov<kBcast15, 0xa>(x, id));
  x = min(x, mov<kBcast31, 0xc>(x, id));
  return unsigned(__builtin_amdgcn_readlane(int(x), 63));
}

// lane l's value, uniform (two SGPRs for a double)
__device__ __forceinline__ double rl(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane(int(b & 0xffffffff), l);
  const int hi = __builtin_amdgcn_readlane(int(b >> 32), l);
  return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}
__device__ __forceinline__ float rl(float v, int l) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l)); }

// (hi, lo) ordered like |v| (the sign bit cleared; bit 31 of hi stays free
// for a candidate flag)
__device__ __forceinline__ void abs_key(double v, unsigned& hi, unsigned& lo) {
  const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(v)) & 0x7fffffffffffffffull;
  hi = unsigned(b >> 32);
  lo = unsigned(b & 0xffffffffu);
}
__device__ __forceinline__ void abs_key(float v, unsigned& hi, unsigned& lo) {
  hi = unsigned(__float_as_int(v)) & 0x7fffffffu;
  lo = 0;
}

// LAPACK idamax over the wave: among the lanes with `cand`, the largest key,
// ties to the smallest `pos`.  Returns the winning lane (uniform); `zero` is
// set when the largest candidate magnitude is exactly zero.
// LO = false: the low key word is zero on every lane (FP32 keys) and is not
// reduced.
template <bool LO = true>
__device__ __forceinline__ int argmax_pos(bool cand, unsigned khi, unsigned klo, int pos, int lane, bool& zero) {
  khi = cand ? (khi | 0x80000000u) : 0u;
  klo = cand ? klo : 0u;
  const unsigned hmax = max_u32(khi);
  const unsigned lmax = LO ? max_u32(khi == hmax ? klo : 0u) : 0u;
  const unsigned long long tie = __ballot(khi == hmax && klo == lmax);
  int who;
  if (__popcll(tie) == 1) {
    who = __ffsll(static_cast<long long>(tie)) - 1;
  } else {
    const bool mine = (tie >> lane) & 1ull;
    const unsigned pm = min_u32(mine ? unsigned(pos) : 0xffffffffu);
    who = __ffsll(static_cast<long long>(__ballot(mine && unsigned(pos) == pm))) - 1;
  }
  zero = hmax == 0x80000000u && lmax == 0u;
  return __builtin_amdgcn_readfirstlane(who);
}

}  // namespace wave
}  // namespace bcr
