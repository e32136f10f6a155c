// Shared device helpers and host-side error plumbing of libbpl_amd (gfx950 / CDNA4).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdarg>
#include <cstdio>

#include "../../include/bpl_amd.h"

namespace bpl {

// message slot / pass direction ids (slot s at pixel p: message received from side s)
enum Dir : int { LEFT = 0, RIGHT = 1, UP = 2, DOWN = 3 };

constexpr int kWave = 64;
constexpr uint8_t kFromTrunc = 255;  // argmin marker: value came from the truncation term

// ------------------------------------------------------------------ host-side errors
void set_error(const char *fmt, ...);
int check_hip(hipError_t e, const char *what);

#define BPL_REQUIRE(cond, code, ...)          \
  do {                                        \
    if (!(cond)) {                            \
      ::bpl::set_error(__VA_ARGS__);          \
      return (code);                          \
    }                                         \
  } while (0)

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

// ------------------------------------------------------------------ lane exchange
// DPP controls (GFX9 encoding): quad_perm [1,0,3,2] / [2,3,0,1], row_half_mirror, row_mirror.
constexpr int kDppXor1 = 0xB1;
constexpr int kDppXor2 = 0x4E;
constexpr int kDppHalfMirror = 0x141;
constexpr int kDppMirror = 0x140;
// ds_swizzle bit-mask mode: and=0x1f, or=0, xor=0x10 -> lane ^ 16 inside 32-lane halves
constexpr int kSwzXor16 = 0x401F;

template <int CTRL>
__device__ __forceinline__ int dpp_i(int v) {
  return __builtin_amdgcn_update_dpp(v, v, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
  return __int_as_float(dpp_i<CTRL>(__float_as_int(v)));
}
__device__ __forceinline__ int swz16_i(int v) { return __builtin_amdgcn_ds_swizzle(v, kSwzXor16); }
__device__ __forceinline__ float swz16_f(float v) { return __int_as_float(swz16_i(__float_as_int(v))); }
__device__ __forceinline__ float readlane_f(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// All-reduce over aligned segments of SEG lanes (16, 32 or 64); every lane of a segment
// receives the same value (the combination tree is symmetric, so float sums agree bitwise).
struct OpSum {
  __device__ __forceinline__ float operator()(float a, float b) const { return a + b; }
};
struct OpMax {
  __device__ __forceinline__ float operator()(float a, float b) const { return fmaxf(a, b); }
};
struct OpMin {
  __device__ __forceinline__ float operator()(float a, float b) const { return fminf(a, b); }
};

template <int SEG, class Op>
__device__ __forceinline__ float seg_reduce(float v, Op op) {
  static_assert(SEG == 16 || SEG == 32 || SEG == 64, "segment width");
  v = op(v, dpp_f<kDppXor1>(v));
  v = op(v, dpp_f<kDppXor2>(v));
  v = op(v, dpp_f<kDppHalfMirror>(v));
  v = op(v, dpp_f<kDppMirror>(v));
  if constexpr (SEG >= 32) v = op(v, swz16_f(v));
  if constexpr (SEG == 64) v = op(readlane_f(v, 0), readlane_f(v, 32));
  return v;
}

// Lexicographic (min value, max key) all-reduce over a segment.
__device__ __forceinline__ void argmin_step(float &v, int &k, float ov, int ok) {
  const bool t = (ov < v) || (ov == v && ok > k);
  v = t ? ov : v;
  k = t ? ok : k;
}

template <int SEG>
__device__ __forceinline__ void seg_argmin(float &v, int &k) {
  argmin_step(v, k, dpp_f<kDppXor1>(v), dpp_i<kDppXor1>(k));
  argmin_step(v, k, dpp_f<kDppXor2>(v), dpp_i<kDppXor2>(k));
  argmin_step(v, k, dpp_f<kDppHalfMirror>(v), dpp_i<kDppHalfMirror>(k));
  argmin_step(v, k, dpp_f<kDppMirror>(v), dpp_i<kDppMirror>(k));
  if constexpr (SEG >= 32) argmin_step(v, k, swz16_f(v), swz16_i(k));
  if constexpr (SEG == 64) {
    float v0 = readlane_f(v, 0), v1 = readlane_f(v, 32);
    int k0 = __builtin_amdgcn_readlane(k, 0), k1 = __builtin_amdgcn_readlane(k, 32);
    argmin_step(v0, k0, v1, k1);
    v = v0;
    k = k0;
  }
}

// Tie key of the reference's shared-memory tree argmin: among equal minima the label
// whose index, bit-reversed over ceil(log2 C) bits, is largest wins.
__device__ __forceinline__ int tie_key(int label, int bits) {
  return bits == 0 ? 0 : (int)(__builtin_bitreverse32((uint32_t)label) >> (32 - bits));
}
inline int tie_bits(int C) {
  int b = 0;
  while ((1 << b) < C) ++b;
  return b;
}

// exp / log from the hardware base-2 instructions with a change of base.
__device__ __forceinline__ float fast_exp(float x) { return __builtin_amdgcn_exp2f(x * 1.4426950408889634f); }
__device__ __forceinline__ float fast_log(float x) { return __builtin_amdgcn_logf(x) * 0.6931471805599453f; }

}  // namespace bpl
