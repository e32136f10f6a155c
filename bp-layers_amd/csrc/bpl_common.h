// Shared device helpers and host-side error plumbing of libbpl_amd (gfx950 / CDNA4).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdarg>
#include <cstdio>
#include <cstdlib>

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

// Scope guard of the timing instrumentation: while bpl_timing_enable(1) is in effect the
// launches made inside the scope are bracketed by events on `stream` and booked to `id`.
class KernelTimer {
 public:
  KernelTimer(int id, hipStream_t stream);
  ~KernelTimer();
  KernelTimer(const KernelTimer &) = delete;
  KernelTimer &operator=(const KernelTimer &) = delete;

 private:
  int id_;
  hipStream_t stream_;
  hipEvent_t start_ = nullptr, stop_ = nullptr;
};

// BPL_LBP_GENERIC=1 selects the generic LBP kernels for every variant (used by tests to
// compare the stereo fast paths with them).
inline bool generic_only() {
  const char *e = getenv("BPL_LBP_GENERIC");
  return e && e[0] == '1';
}
// BPL_LBP_NO_P1P2=1 turns the P1/P2 sweep of the stereo forward off (its scanlines take the
// 7-candidate fast step), for comparisons of the two.
inline bool p1p2_off() {
  const char *e = getenv("BPL_LBP_NO_P1P2");
  return e && e[0] == '1';
}

// ------------------------------------------------------------------ lane exchange
// DPP controls (GFX9 encoding): quad_perm [1,0,3,2] / [2,3,0,1], row_half_mirror, row_mirror.
constexpr int kDppXor1 = 0xB1;
constexpr int kDppXor2 = 0x4E;
constexpr int kDppHalfMirror = 0x141;
constexpr int kDppMirror = 0x140;

template <int CTRL>
__device__ __forceinline__ int dpp_i(int v) {
  return __builtin_amdgcn_update_dpp(v, v, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
  return __int_as_float(dpp_i<CTRL>(__float_as_int(v)));
}
__device__ __forceinline__ float readlane_f(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// gfx950 row/half swaps with both operands = v: `lo` receives the even 16-lane row (or the
// lower 32-lane half) of each pair, `hi` the odd row (upper half), in every lane of the pair.
__device__ __forceinline__ void swap_rows(int v, int &lo, int &hi) {
  const auto r = __builtin_amdgcn_permlane16_swap((unsigned)v, (unsigned)v, false, false);
  lo = (int)r[0];
  hi = (int)r[1];
}
__device__ __forceinline__ void swap_halves(int v, int &lo, int &hi) {
  const auto r = __builtin_amdgcn_permlane32_swap((unsigned)v, (unsigned)v, false, false);
  lo = (int)r[0];
  hi = (int)r[1];
}
__device__ __forceinline__ void swap_rows(float v, float &lo, float &hi) {
  int a, b;
  swap_rows(__float_as_int(v), a, b);
  lo = __int_as_float(a);
  hi = __int_as_float(b);
}
__device__ __forceinline__ void swap_halves(float v, float &lo, float &hi) {
  int a, b;
  swap_halves(__float_as_int(v), a, b);
  lo = __int_as_float(a);
  hi = __int_as_float(b);
}

// Order-preserving integer key of a float (signed compare of keys == float compare, with
// -0 < +0); key_to_float inverts it bit for bit.
__device__ __forceinline__ int float_key(float f) {
  const int b = __float_as_int(f);
  return b ^ ((b >> 31) & 0x7FFFFFFF);
}
__device__ __forceinline__ float key_to_float(int k) { return __int_as_float(k ^ ((k >> 31) & 0x7FFFFFFF)); }

// Full-wave all-reduces whose row stages fold the DPP move into the arithmetic instruction
// (bound_ctrl set: every lane of these patterns has a valid source).
template <int CTRL>
__device__ __forceinline__ int dpp_src_i(int v) {
  return __builtin_amdgcn_mov_dpp(v, CTRL, 0xF, 0xF, true);
}
template <int CTRL>
__device__ __forceinline__ float dpp_src_f(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, true));
}
__device__ __forceinline__ int wave_min_i(int v) {
  v = min(v, dpp_src_i<kDppXor1>(v));
  v = min(v, dpp_src_i<kDppXor2>(v));
  v = min(v, dpp_src_i<kDppHalfMirror>(v));
  v = min(v, dpp_src_i<kDppMirror>(v));
  int lo, hi;
  swap_rows(v, lo, hi);
  v = min(lo, hi);
  swap_halves(v, lo, hi);
  return min(lo, hi);
}
__device__ __forceinline__ int wave_max_i(int v) {
  v = max(v, dpp_src_i<kDppXor1>(v));
  v = max(v, dpp_src_i<kDppXor2>(v));
  v = max(v, dpp_src_i<kDppHalfMirror>(v));
  v = max(v, dpp_src_i<kDppMirror>(v));
  int lo, hi;
  swap_rows(v, lo, hi);
  v = max(lo, hi);
  swap_halves(v, lo, hi);
  return max(lo, hi);
}
// Float minimum over the wave.  Built with -fno-honor-nans (the LBP translation units) the
// DPP moves fold into v_min_f32_dpp; otherwise the compiler adds canonicalisations (same
// values).
__device__ __forceinline__ float wave_min_f(float v) {
  v = fminf(v, dpp_src_f<kDppXor1>(v));
  v = fminf(v, dpp_src_f<kDppXor2>(v));
  v = fminf(v, dpp_src_f<kDppHalfMirror>(v));
  v = fminf(v, dpp_src_f<kDppMirror>(v));
  float lo, hi;
  swap_rows(v, lo, hi);
  v = fminf(lo, hi);
  swap_halves(v, lo, hi);
  return fminf(lo, hi);
}
// Workgroup b of a launch runs on XCD b mod 8 (the dispatcher deals workgroups to the eight
// XCDs in turn).  This renumbering gives every XCD a contiguous range of logical indices, so
// that neighbouring scanlines of a vertical phase, which share cache lines -- 4-byte accesses
// to per-pixel planes (pairwise parameters, their gradients), and label rows that are not a
// whole number of lines (19 or 25 labels: 76 or 100 bytes) -- run on one XCD and meet in
// its L2 instead of each XCD reading and writing back its own partial lines.
__device__ __forceinline__ int xcd_contiguous_block(int b, int nblocks) {
  constexpr int kXcd = 8;
  const int k = b % kXcd, j = b / kXcd;
  const int q = nblocks / kXcd, r = nblocks % kXcd;
  return k * q + min(k, r) + j;
}
__device__ __forceinline__ float wave_sum_f(float v) {
  v = v + dpp_src_f<kDppXor1>(v);
  v = v + dpp_src_f<kDppXor2>(v);
  v = v + dpp_src_f<kDppHalfMirror>(v);
  v = v + dpp_src_f<kDppMirror>(v);
  float lo, hi;
  swap_rows(v, lo, hi);
  v = lo + hi;
  swap_halves(v, lo, hi);
  return lo + hi;
}

// Reduce-scatter of eight per-lane partial sums v[0..7] over the wave.  The two-operand swaps
// exchange the upper 32 lanes of the first operand with the lower 32 of the second
// (v_permlane32_swap) and the odd 16-lane rows of the first with the even rows of the second
// (v_permlane16_swap), so each swap-and-add halves the lanes a sum spans while keeping two
// sums in one register.  Result: the total of v[b] is held by every lane of 16-lane row
// b >> 1 in the register returned for b & 1.
__device__ __forceinline__ void swap_halves2(float a, float b, float &x, float &y) {
  const auto r = __builtin_amdgcn_permlane32_swap((unsigned)__float_as_int(a), (unsigned)__float_as_int(b), false, false);
  x = __int_as_float((int)r[0]);
  y = __int_as_float((int)r[1]);
}
__device__ __forceinline__ void swap_rows2(float a, float b, float &x, float &y) {
  const auto r = __builtin_amdgcn_permlane16_swap((unsigned)__float_as_int(a), (unsigned)__float_as_int(b), false, false);
  x = __int_as_float((int)r[0]);
  y = __int_as_float((int)r[1]);
}
__device__ __forceinline__ float row_sum_f(float v) {
  v = v + dpp_src_f<kDppXor1>(v);
  v = v + dpp_src_f<kDppXor2>(v);
  v = v + dpp_src_f<kDppHalfMirror>(v);
  return v + dpp_src_f<kDppMirror>(v);
}
__device__ __forceinline__ void reduce_scatter8(const float *v, float &even, float &odd) {
  float u[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {  // lanes 0-31: v[j] over both halves, lanes 32-63: v[j + 4]
    float x, y;
    swap_halves2(v[j], v[j + 4], x, y);
    u[j] = x + y;
  }
  float z[2];
#pragma unroll
  for (int j = 0; j < 2; ++j) {  // rows 0..3: u[j].lo, u[j+2].lo, u[j].hi, u[j+2].hi
    float x, y;
    swap_rows2(u[j], u[j + 2], x, y);
    z[j] = x + y;
  }
  even = row_sum_f(z[0]);  // rows 0..3 hold v[0], v[2], v[4], v[6]
  odd = row_sum_f(z[1]);   // rows 0..3 hold v[1], v[3], v[5], v[7]
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
  v = op(v, dpp_src_f<kDppXor1>(v));  // the moves fold into the operation (v_add_f32_dpp, ...)
  v = op(v, dpp_src_f<kDppXor2>(v));
  v = op(v, dpp_src_f<kDppHalfMirror>(v));
  v = op(v, dpp_src_f<kDppMirror>(v));
  if constexpr (SEG >= 32) {
    float lo, hi;
    swap_rows(v, lo, hi);
    v = op(lo, hi);
  }
  if constexpr (SEG == 64) {
    float lo, hi;
    swap_halves(v, lo, hi);
    v = op(lo, hi);
  }
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
  if constexpr (SEG >= 32) {
    float v0, v1;
    int k0, k1;
    swap_rows(v, v0, v1);
    swap_rows(k, k0, k1);
    argmin_step(v0, k0, v1, k1);
    v = v0;
    k = k0;
  }
  if constexpr (SEG == 64) {
    float v0, v1;
    int k0, k1;
    swap_halves(v, v0, v1);
    swap_halves(k, k0, k1);
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

// Workgroup barrier that orders LDS accesses only: the fences are restricted to the local
// address space, so the barrier waits for outstanding LDS operations (lgkmcnt) but not for
// global loads in flight (a plain __syncthreads would also drain the prefetch ring).
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Orders the LDS accesses of one wave: writes by its lanes are visible to its later reads
// (no other wave is waited for).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

// Uniform (SGPR) copy of a wave-uniform 64-bit pointer: a buffer resource built from it is
// then known to be uniform (otherwise every buffer access becomes a readfirstlane loop).
__device__ __forceinline__ void *uniform_ptr(const void *p) {
  const uint64_t v = (uint64_t)(uintptr_t)p;
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return (void *)(uintptr_t)(((uint64_t)hi << 32) | lo);
}

// Raw buffer access: a resource over `bytes` bytes from a wave-uniform base, a per-lane
// byte offset `voff` and a wave-uniform byte offset `soff` (a scanline's row).  The
// address is formed by the memory unit, so the kernel carries no per-lane 64-bit address;
// an access whose offsets reach past `bytes` reads 0 and a store there is dropped
// (kBufDrop as `voff` drops a lane).
constexpr uint32_t kBufDrop = 0x7FFFFFFFu;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void *base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(uniform_ptr(base), (short)0, (int)bytes, 0x00020000);
}
// AUX: the access's cache-policy bits (0: default; 2: non-temporal)
template <int AUX = 0>
__device__ __forceinline__ float buf_ld(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, AUX));
}
template <int AUX = 0>
__device__ __forceinline__ void buf_st(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff, float v) {
  __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, voff, soff, AUX);
}
__device__ __forceinline__ void buf_st8(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff, uint8_t v) {
  __builtin_amdgcn_raw_buffer_store_b8(v, r, voff, soff, 0);
}
// LPL consecutive floats per lane (one dword / dwordx2 / dwordx3 / dwordx4 access) at
// base + voff + soff; voff must be a multiple of 4.
template <int LPL, int AUX = 0>
__device__ __forceinline__ void buf_ld_row(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff, float *d) {
  static_assert(LPL >= 1 && LPL <= 4, "1 to 4 floats per lane");
  if constexpr (LPL == 1) {
    d[0] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, AUX));
  } else if constexpr (LPL == 2) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, AUX);
    d[0] = __uint_as_float(v[0]); d[1] = __uint_as_float(v[1]);
  } else if constexpr (LPL == 3) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b96(r, voff, soff, AUX);
    d[0] = __uint_as_float(v[0]); d[1] = __uint_as_float(v[1]); d[2] = __uint_as_float(v[2]);
  } else {
    const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, AUX);
    d[0] = __uint_as_float(v[0]); d[1] = __uint_as_float(v[1]);
    d[2] = __uint_as_float(v[2]); d[3] = __uint_as_float(v[3]);
  }
}
template <int LPL, int AUX = 0>
__device__ __forceinline__ void buf_st_row(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff, const float *s) {
  static_assert(LPL >= 1 && LPL <= 4, "1 to 4 floats per lane");
  if constexpr (LPL == 1) {
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(s[0]), r, voff, soff, AUX);
  } else if constexpr (LPL == 2) {
    decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0)) v;
    v[0] = __float_as_uint(s[0]); v[1] = __float_as_uint(s[1]);
    __builtin_amdgcn_raw_buffer_store_b64(v, r, voff, soff, AUX);
  } else if constexpr (LPL == 3) {
    decltype(__builtin_amdgcn_raw_buffer_load_b96(r, 0, 0, 0)) v;
    v[0] = __float_as_uint(s[0]); v[1] = __float_as_uint(s[1]); v[2] = __float_as_uint(s[2]);
    __builtin_amdgcn_raw_buffer_store_b96(v, r, voff, soff, AUX);
  } else {
    decltype(__builtin_amdgcn_raw_buffer_load_b128(r, 0, 0, 0)) v;
    v[0] = __float_as_uint(s[0]); v[1] = __float_as_uint(s[1]);
    v[2] = __float_as_uint(s[2]); v[3] = __float_as_uint(s[3]);
    __builtin_amdgcn_raw_buffer_store_b128(v, r, voff, soff, AUX);
  }
}

// exp / log from the hardware base-2 instructions with a change of base.
__device__ __forceinline__ float fast_exp(float x) { return __builtin_amdgcn_exp2f(x * 1.4426950408889634f); }
__device__ __forceinline__ float fast_log(float x) { return __builtin_amdgcn_logf(x) * 0.6931471805599453f; }

}  // namespace bpl
