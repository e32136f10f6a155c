// SAD matching-cost kernels on gfx950: stereo cost volume and fused flow window SAD with
// min-projections, forward and (deterministic, gather-based) backward.
//
// Semantics follow ops/sad/src/stereo_sad_kernel.cu:16-108 and
// ops/flow_mp_sad/src/flow_mp_sad_kernel.cu:10-200 (fill value 40 outside the image,
// sign(diff) with a 1e-15 dead zone in the backward, first-index argmins in the flow
// min-projections); every batch element is computed.
#include "bpl_common.h"

namespace bpl {
namespace {

constexpr float kFill = 40.f;
constexpr float kSignEps = 1e-15f;

__device__ __forceinline__ float sgn_eps(float t) { return fabsf(t) > kSignEps ? (t > 0.f ? 1.f : -1.f) : 0.f; }

// ----------------------------------------------------------------------------- stereo fwd
// Block: 64 pixels of one row x 64 disparities; thread = 4 pixels x 4 disparities.  Rows of
// f0 and f1 are staged channel-chunk by channel-chunk in LDS; each f1 value is reused by up
// to 16 (pixel, disparity) pairs of a thread.
constexpr int ST_TX = 64, ST_TD = 64, ST_CB = 16;

__global__ __launch_bounds__(256) void sad_stereo_fwd_k(const float *__restrict__ f0, const float *__restrict__ f1,
                                                        float *__restrict__ cost, int C, int H, int W, int D,
                                                        int mind, int dpasses) {
  __shared__ float s0[ST_CB][ST_TX];
  __shared__ float s1[ST_CB][ST_TX + ST_TD];
  const int x0 = blockIdx.x * ST_TX;
  const int y = blockIdx.y;
  const int n = blockIdx.z / dpasses;
  const int dp = blockIdx.z - n * dpasses;
  const int d0 = mind + dp * ST_TD;
  const int t = threadIdx.x;
  const int dg = t & 15, xg = t >> 4;
  const int xs = x0 - d0 - (ST_TD - 1);  // f1 column of s1[.][0]
  const int64_t HW = (int64_t)H * W;
  const float *F0 = f0 + (int64_t)n * C * HW + (int64_t)y * W;
  const float *F1 = f1 + (int64_t)n * C * HW + (int64_t)y * W;
  float acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = 0.f;

  for (int c0 = 0; c0 < C; c0 += ST_CB) {
    for (int i = t; i < ST_CB * ST_TX; i += 256) {
      const int cc = i / ST_TX, xx = i - cc * ST_TX;
      const int c = c0 + cc, xc = x0 + xx;
      s0[cc][xx] = (c < C && xc < W) ? F0[(int64_t)c * HW + xc] : 0.f;
    }
    for (int i = t; i < ST_CB * (ST_TX + ST_TD); i += 256) {
      const int cc = i / (ST_TX + ST_TD), j = i - cc * (ST_TX + ST_TD);
      const int c = c0 + cc, xc = xs + j;
      s1[cc][j] = (c < C && xc >= 0 && xc < W) ? F1[(int64_t)c * HW + xc] : 0.f;
    }
    __syncthreads();
    const int nc = min(ST_CB, C - c0);
    const int jb = xg * 4 - dg * 4 + ST_TD - 4;
    for (int cc = 0; cc < nc; ++cc) {
      float a4[4], v[7];
#pragma unroll
      for (int a = 0; a < 4; ++a) a4[a] = s0[cc][xg * 4 + a];
#pragma unroll
      for (int j = 0; j < 7; ++j) v[j] = s1[cc][jb + j];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] += fabsf(a4[a] - v[a - b + 3]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int x = x0 + xg * 4 + a;
    if (x >= W) continue;
    float *out = cost + (((int64_t)n * H + y) * W + x) * D;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int i = dp * ST_TD + dg * 4 + b;
      const int d = mind + i;
      if (i < D) out[i] = (x - d >= 0 && x - d < W) ? acc[a][b] : kFill;
    }
  }
}

// Stereo SAD with the matching softmax fused in: y = sign * softmax(scale * SAD) over the D
// disparities of each pixel (matching.py:22-25 and :45-47 with corenet.py:97-104: scale =
// -1/T; sign = -1 gives the message-passing cost -prob, message_passing_op_cuda.py:162).
// Block: 64 pixels of one row x all D disparities.  The disparity passes of sad_stereo_fwd_k
// run one after another in the block with the same per-thread arithmetic and leave their SAD
// tile in the block's own [64][D] slice of y, which stays in L2 (a tile in LDS instead costs
// the SAD loop its occupancy: 1.18 vs 0.60 ms at 384x1248x192x64); then each wave normalises
// 16 of the pixels with the arithmetic of bpl_softmax_labels (label c on lane c mod 64, LPL
// labels per lane, the same max and sum trees) and overwrites the slice, so y equals
// bpl_sad_stereo_fwd followed by bpl_softmax_labels bit for bit.
template <int LPL>
__global__ __launch_bounds__(256) void sad_stereo_softmax_k(const float *__restrict__ f0, const float *__restrict__ f1,
                                                            float *y, int C, int H, int W, int D, int mind,
                                                            float scale, float sign) {
  __shared__ float s0[ST_CB][ST_TX];
  __shared__ float s1[ST_CB][ST_TX + ST_TD];
  const int x0 = blockIdx.x * ST_TX;
  const int row = blockIdx.y;
  const int n = blockIdx.z;
  const int t = threadIdx.x;
  const int dg = t & 15, xg = t >> 4;
  const int64_t HW = (int64_t)H * W;
  const float *F0 = f0 + (int64_t)n * C * HW + (int64_t)row * W;
  const float *F1 = f1 + (int64_t)n * C * HW + (int64_t)row * W;
  float *sv = y + (((int64_t)n * H + row) * W + x0) * D;  // the block's [64][D] slice of y
  const int dpasses = (D + ST_TD - 1) / ST_TD;
  for (int dp = 0; dp < dpasses; ++dp) {
    const int d0 = mind + dp * ST_TD;
    const int xs = x0 - d0 - (ST_TD - 1);  // f1 column of s1[.][0]
    float acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] = 0.f;
    for (int c0 = 0; c0 < C; c0 += ST_CB) {
      for (int i = t; i < ST_CB * ST_TX; i += 256) {
        const int cc = i / ST_TX, xx = i - cc * ST_TX;
        const int c = c0 + cc, xc = x0 + xx;
        s0[cc][xx] = (c < C && xc < W) ? F0[(int64_t)c * HW + xc] : 0.f;
      }
      for (int i = t; i < ST_CB * (ST_TX + ST_TD); i += 256) {
        const int cc = i / (ST_TX + ST_TD), j = i - cc * (ST_TX + ST_TD);
        const int c = c0 + cc, xc = xs + j;
        s1[cc][j] = (c < C && xc >= 0 && xc < W) ? F1[(int64_t)c * HW + xc] : 0.f;
      }
      __syncthreads();
      const int nc = min(ST_CB, C - c0);
      const int jb = xg * 4 - dg * 4 + ST_TD - 4;
      for (int cc = 0; cc < nc; ++cc) {
        float a4[4], v[7];
#pragma unroll
        for (int a = 0; a < 4; ++a) a4[a] = s0[cc][xg * 4 + a];
#pragma unroll
        for (int j = 0; j < 7; ++j) v[j] = s1[cc][jb + j];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
          for (int b = 0; b < 4; ++b) acc[a][b] += fabsf(a4[a] - v[a - b + 3]);
      }
      __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const int px = xg * 4 + a, x = x0 + px;
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const int i = dp * ST_TD + dg * 4 + b;
        const int d = mind + i;
        if (i < D && x < W) sv[px * D + i] = (x - d >= 0 && x - d < W) ? acc[a][b] : kFill;
      }
    }
  }
  // the tile written by all waves of the block is read back by other waves
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __syncthreads();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  const int lane = t & 63, w = t >> 6;
  for (int q = 0; q < ST_TX / 4; ++q) {
    const int px = q * 4 + w, x = x0 + px;
    if (x >= W) break;  // wave-uniform; later q only go further right
    const float *xr = sv + px * D;
    float v[LPL];
    float mx = -__builtin_inff();
#pragma unroll
    for (int k = 0; k < LPL; ++k) {
      const int c = lane + 64 * k;
      v[k] = c < D ? xr[c] * scale : -__builtin_inff();
      mx = fmaxf(mx, v[k]);
    }
    mx = seg_reduce<64>(mx, OpMax());
    float sum = 0.f;
#pragma unroll
    for (int k = 0; k < LPL; ++k) {
      const int c = lane + 64 * k;
      v[k] = c < D ? fast_exp(v[k] - mx) : 0.f;
      sum += v[k];
    }
    sum = seg_reduce<64>(sum, OpSum());
    const float r = sign / sum;
    float *yr = sv + px * D;
#pragma unroll
    for (int k = 0; k < LPL; ++k) {
      const int c = lane + 64 * k;
      if (c < D) yr[c] = v[k] * r;
    }
  }
}

// general step: one thread per output, d = int(i * step + min_disp) (stereo_sad_kernel.cu:35)
__global__ void sad_stereo_fwd_step_k(const float *__restrict__ f0, const float *__restrict__ f1, float *__restrict__ cost,
                                      int N, int C, int H, int W, int D, int mind, float step) {
  const int64_t total = (int64_t)N * H * W * D;
  const int64_t HW = (int64_t)H * W;
  for (int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; o < total; o += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(o % D);
    const int64_t p = o / D;
    const int x = (int)(p % W);
    const int y = (int)((p / W) % H);
    const int n = (int)(p / HW);
    const int d = (int)((float)i * step + (float)mind);
    float r = kFill;
    if (x - d >= 0 && x - d < W) {
      r = 0.f;
      for (int c = 0; c < C; ++c) {
        const int64_t b = ((int64_t)n * C + c) * HW + (int64_t)y * W;
        r += fabsf(f0[b + x] - f1[b + x - d]);
      }
    }
    cost[o] = r;
  }
}

// ----------------------------------------------------------------------------- stereo bwd
// Block: one row, a tile of 32 pixels, 256 threads = 16 pixel pairs x 16 channel lanes.  A
// thread accumulates its 2 pixels for the channels cl, cl + 16, cl + 32, cl + 48 of a group of
// 64 channels.  Disparities are taken in chunks of SB_DC; per chunk the gradient tiles
//   G0[k][x] = g[y, x, k]      (0 where x - d is outside the image)   -> df0 gather
//   G1[k][x] = g[y, x + d, k]  (0 where x + d is outside the image)   -> df1 gather
// are staged once (the diagonal G1 row by row of its source pixel, so the global reads stay
// contiguous), then, for each 16-channel part of the group, the f1 segment
// [x0 - d_hi, x0 + 31 - d_lo] and the f0 segment [x0 + d_lo, x0 + 31 + d_hi] of the chunk.  Each
// thread sums its terms in disparity order (the reference's order) from LDS, reading one new f
// value per disparity per side; the invalid pairs carry g = 0, so the loop has no range tests.
// Chunking keeps the block's LDS at 29 KB, so five blocks (20 waves) share a CU.
constexpr int SB_TX = 32, SB_CB = 16, SB_GP = SB_TX + 1, SB_DC = 64, SB_CG = 4;
constexpr int SB_SPAN = SB_TX + SB_DC - 1;

// acc + sign(t) * g with the reference's dead zone |t| <= 1e-15 -> 0, in two VALU
// instructions after the difference t.  The indicator [|t| > eps] is one clamped FMA:
// 2^127 |t| - 2^127 eps has the sign of |t| - eps exactly (power-of-two scaling, one
// rounding), and when positive it is at least 2^127 ulp(eps) = 2^54, so the clamp to [0, 1]
// yields exactly 0 or 1 (NaN t clamps to 0).  The sign of t is a sign-bit flip of g (exact),
// and fma(1, ±g, acc) = acc ± g, fma(0, ±g, acc) = acc for finite g: the sums are those of
// adding the selected term.  NEG subtracts the term instead.
template <bool NEG>
__device__ __forceinline__ float acc_sgn(float acc, float t, float g) {
  constexpr float kScale = 0x1p127f, kScaledEps = kSignEps * 0x1p127f;
  const float on = __builtin_amdgcn_fmed3f(fmaf(fabsf(t), kScale, -kScaledEps), 0.f, 1.f);
  const float v = __int_as_float(__float_as_int(g) ^ (__float_as_int(t) & (int)0x80000000));
  return fmaf(NEG ? -on : on, v, acc);
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void sad_stereo_bwd_k(const float *__restrict__ f0, const float *__restrict__ f1,
                                                        const float *__restrict__ g, float *__restrict__ df0,
                                                        float *__restrict__ df1, int C, int H, int W, int D, int mind) {
  __shared__ float G0[SB_DC * SB_GP];      // [k - d0][x]
  __shared__ float G1[SB_DC * SB_GP];      // [k - d0][x]
  __shared__ float F1[SB_CB * SB_SPAN];    // [channel][j]: f1 at x0 - (mind + d0 + SB_DC - 1) + j
  __shared__ float F0[SB_CB * SB_SPAN];    // [channel][j]: f0 at x0 + mind + d0 + j
  // rows go to the XCDs in contiguous bands: the f0 and f1 columns that neighbouring tiles of a
  // row both stage are read through one L2
  const int gx = (int)gridDim.x;
  const int tb = xcd_contiguous_block((int)blockIdx.x + gx * (int)blockIdx.y, gx * (int)gridDim.y);
  const int x0 = (tb % gx) * SB_TX;
  const int y = tb / gx;
  const int n = blockIdx.z;
  const int t = threadIdx.x;
  const int64_t HW = (int64_t)H * W;
  const float *Gr = g + ((int64_t)n * H + y) * (int64_t)W * D;
  const int xg = t & 15;
  const int cl = t >> 4;
  const int xa = x0 + 2 * xg;
  for (int cg0 = 0; cg0 < C; cg0 += SB_CB * SB_CG) {
    float r00[SB_CG], r01[SB_CG], r10[SB_CG], r11[SB_CG], a0[SB_CG], a1[SB_CG], b0[SB_CG], b1[SB_CG];
#pragma unroll
    for (int p = 0; p < SB_CG; ++p) {
      const int c = cg0 + p * SB_CB + cl;
      const int64_t b = ((int64_t)n * C + (c < C ? c : C - 1)) * HW + (int64_t)y * W;
      a0[p] = xa < W ? f0[b + xa] : 0.f;
      a1[p] = xa + 1 < W ? f0[b + xa + 1] : 0.f;
      b0[p] = xa < W ? f1[b + xa] : 0.f;
      b1[p] = xa + 1 < W ? f1[b + xa + 1] : 0.f;
      r00[p] = r01[p] = r10[p] = r11[p] = 0.f;
    }
    for (int d0 = 0; d0 < D; d0 += SB_DC) {
      const int dn = min(SB_DC, D - d0);
      __syncthreads();  // the previous chunk's readers are done
      // compile-time trip counts and branch-free bodies: every load reads a clamped, in-bounds
      // address and a select stores 0 for the invalid entries, so the unrolled loads are in
      // flight together instead of one per branch
#pragma unroll 4
      for (int i = t; i < SB_TX * SB_DC; i += 256) {
        const int xl = i / SB_DC, kk = i - xl * SB_DC;
        const int x = x0 + xl, xm = x - (mind + d0 + kk);
        const bool ok = kk < dn && x < W && xm >= 0 && xm < W;
        const float v = Gr[(int64_t)min(x, W - 1) * D + d0 + min(kk, dn - 1)];
        G0[kk * SB_GP + xl] = ok ? v : 0.f;
      }
      // source pixel p = x0 + mind + m carries the entries k = m - xl of the tile's pixels x0 + xl
#pragma unroll 4
      for (int i = t; i < (SB_DC + SB_TX - 1) * SB_TX; i += 256) {
        const int mm = i / SB_TX, xl = i - mm * SB_TX;
        const int kk = mm - xl;  // k - d0
        const int x = x0 + xl, pp = x0 + mind + d0 + mm;
        const bool in_chunk = kk >= 0 && kk < dn;
        const bool ok = in_chunk && x < W && pp >= 0 && pp < W;
        const float v = Gr[(int64_t)min(max(pp, 0), W - 1) * D + d0 + min(max(kk, 0), dn - 1)];
        if (in_chunk) G1[kk * SB_GP + xl] = ok ? v : 0.f;
      }
#pragma unroll
      for (int p = 0; p < SB_CG; ++p) {
        const int cb = cg0 + p * SB_CB;
        if (cb >= C) break;
        if (p > 0) __syncthreads();  // the previous part's readers of F are done
#pragma unroll 3
        for (int i = t; i < SB_CB * SB_SPAN; i += 256) {
          const int cc = i / SB_SPAN, j = i - cc * SB_SPAN;
          const int c = cb + cc;
          const int64_t b = ((int64_t)n * C + min(c, C - 1)) * HW + (int64_t)y * W;
          const int xs1 = x0 - (mind + d0 + SB_DC - 1) + j, xs0 = x0 + mind + d0 + j;
          const float v1 = f1[b + min(max(xs1, 0), W - 1)], v0 = f0[b + min(max(xs0, 0), W - 1)];
          F1[i] = (c < C && xs1 >= 0 && xs1 < W) ? v1 : 0.f;
          F0[i] = (c < C && xs0 >= 0 && xs0 < W) ? v0 : 0.f;
        }
        __syncthreads();
        // f1[x - d] = F1c[-k], f0[x + d] = F0c[k] for pixel xa (k = d - mind, absolute)
        const float *F1c = F1 + cl * SB_SPAN + 2 * xg + d0 + SB_DC - 1;
        const float *F0c = F0 + cl * SB_SPAN + 2 * xg - d0;
        const float *G0c = G0 + 2 * xg - d0 * SB_GP;
        const float *G1c = G1 + 2 * xg - d0 * SB_GP;
        float s00 = r00[p], s01 = r01[p], s10 = r10[p], s11 = r11[p];
        const float fa0 = a0[p], fa1 = a1[p], fb0 = b0[p], fb1 = b1[p];
        float v1 = F1c[1 - d0];  // f1[xa + 1 - d] at k = d0
        float v0 = F0c[d0];      // f0[xa + d] at k = d0
#pragma unroll 4
        for (int k = d0; k < d0 + dn; ++k) {
          const float v1n = F1c[-k];     // f1[xa - d]; f1[xa + 1 - d] is the previous step's
          const float v0n = F0c[k + 1];  // f0[xa + 1 + d]; f0[xa + d] is the previous step's
          const float g00 = G0c[k * SB_GP], g01 = G0c[k * SB_GP + 1];
          const float g10 = G1c[k * SB_GP], g11 = G1c[k * SB_GP + 1];
          s00 = acc_sgn<false>(s00, fa0 - v1n, g00);
          s01 = acc_sgn<false>(s01, fa1 - v1, g01);
          s10 = acc_sgn<true>(s10, v0 - fb0, g10);
          s11 = acc_sgn<true>(s11, v0n - fb1, g11);
          v1 = v1n;
          v0 = v0n;
        }
        r00[p] = s00; r01[p] = s01; r10[p] = s10; r11[p] = s11;
      }
    }
#pragma unroll
    for (int p = 0; p < SB_CG; ++p) {
      const int c = cg0 + p * SB_CB + cl;
      if (c < C) {
        const int64_t b = ((int64_t)n * C + c) * HW + (int64_t)y * W;
        if (xa < W) { df0[b + xa] = r00[p]; df1[b + xa] = r10[p]; }
        if (xa + 1 < W) { df0[b + xa + 1] = r01[p]; df1[b + xa + 1] = r11[p]; }
      }
    }
  }
}

// ----------------------------------------------------------------------------- flow fwd
// Block: 8 rows x 64 pixels, thread = 2 adjacent pixels.  For each vertical displacement v
// the f1 rows y+ov+v-h of the tile (plus the horizontal halo) are staged in LDS channel chunk
// by chunk; S(u,v) for all u is accumulated in registers, then cv_v[v] (min over u) is
// written and the running cv_u (min over v) updated -- both with first-index ties.
constexpr int FL_TY = 4, FL_TX = 32, FL_CB = 8, FL_G = 4;

// Block: 4 rows x 32 pixels x 4 groups of vertical displacements (512 threads).  Group g owns
// the displacements vi = g, g+4, ...; the 4 groups advance together, so one staging of the f1
// rows for 4 consecutive displacements (and 8 channels) serves all of them.  For each vi a thread
// forms S(u, vi) for all K horizontal u in registers, writes the v-projection (min over u,
// first index) and keeps a running u-projection (min over its vi); the groups' partial
// u-projections are merged through LDS (smaller vi wins ties: the reference's first index).
template <int K>
__global__ __launch_bounds__(512) void sad_flow_fwd_k(const float *__restrict__ f0, const float *__restrict__ f1,
                                                      float *__restrict__ cv_u, float *__restrict__ cv_v,
                                                      float *__restrict__ u_star, float *__restrict__ v_star,
                                                      int C, int H, int W, int sws, int ou, int ov, int bu, int bv) {
  constexpr int G = FL_G;
  constexpr int SW = FL_TX + K - 1;  // staged f1 columns
  constexpr int SR = FL_TY + G - 1;  // staged f1 rows (4 consecutive displacements)
  constexpr int NI = (K + G - 1) / G;
  constexpr int NP = FL_TY * FL_TX;
  // staging tiles, and after the main loop the partial u-projections of groups 1..G-1
  constexpr size_t kStage = sizeof(float) * (FL_CB * SR * SW + FL_CB * FL_TY * FL_TX);
  constexpr size_t kMerge = (sizeof(float) + 1) * (G - 1) * NP * K;
  __shared__ __align__(16) unsigned char lds[kStage > kMerge ? kStage : kMerge];
  auto s1 = reinterpret_cast<float (*)[SR][SW]>(lds);
  auto s0 = reinterpret_cast<float (*)[FL_TY][FL_TX]>(lds + sizeof(float) * FL_CB * SR * SW);
  auto m_v = reinterpret_cast<float (*)[NP][K]>(lds);
  auto m_i = reinterpret_cast<unsigned char (*)[NP][K]>(lds + sizeof(float) * (G - 1) * NP * K);
  const int hs = sws / 2;
  // tiles go to the XCDs in contiguous ranges (bands of tile rows): the f1 rows and halo
  // columns that neighbouring tiles both stage are read through one L2
  const int gx = (int)gridDim.x;
  const int tb = xcd_contiguous_block((int)blockIdx.x + gx * (int)blockIdx.y, gx * (int)gridDim.y);
  const int x0 = (tb % gx) * FL_TX;
  const int y0 = (tb / gx) * FL_TY;
  const int n = blockIdx.z;
  const int t = threadIdx.x;
  const int g = t / (FL_TY * FL_TX);
  const int pl = t - g * (FL_TY * FL_TX);
  const int ty = pl / FL_TX, tx = pl - ty * FL_TX;
  const int y = y0 + ty, x = x0 + tx;
  const int64_t HW = (int64_t)H * W;
  const float *F0 = f0 + (int64_t)n * C * HW;
  const float *F1 = f1 + (int64_t)n * C * HW;
  float bmin[K];
  int barg[K];
#pragma unroll
  for (int u = 0; u < K; ++u) { bmin[u] = 9999999.f; barg[u] = 0; }

  for (int i = 0; i < NI; ++i) {
    const int vi = g + G * i;
    const int dv0 = ov + G * i - hs;  // displacement of group 0 at this round
    float S[K];
#pragma unroll
    for (int u = 0; u < K; ++u) S[u] = 0.f;
    for (int c0 = 0; c0 < C; c0 += FL_CB) {
      __syncthreads();
      for (int e = t; e < FL_CB * SR * SW; e += G * FL_TY * FL_TX) {
        const int cc = e / (SR * SW);
        const int r = e - cc * SR * SW;
        const int ry = r / SW, j = r - ry * SW;
        const int c = c0 + cc, yy = y0 + ry + dv0, xx = x0 + ou - hs + j;
        s1[cc][ry][j] = (c < C && yy >= 0 && yy < H && xx >= 0 && xx < W) ? F1[(int64_t)c * HW + (int64_t)yy * W + xx] : 0.f;
      }
      for (int e = t; e < FL_CB * FL_TY * FL_TX; e += G * FL_TY * FL_TX) {
        const int cc = e / (FL_TY * FL_TX);
        const int r = e - cc * FL_TY * FL_TX;
        const int ry = r / FL_TX, j = r - ry * FL_TX;
        const int c = c0 + cc, yy = y0 + ry, xx = x0 + j;
        s0[cc][ry][j] = (c < C && yy < H && xx < W) ? F0[(int64_t)c * HW + (int64_t)yy * W + xx] : 0.f;
      }
      __syncthreads();
      if (vi < K) {
        const int nc = min(FL_CB, C - c0);
        for (int cc = 0; cc < nc; ++cc) {
          const float a = s0[cc][ty][tx];
          const float *row = &s1[cc][ty + g][tx];
#pragma unroll
          for (int u = 0; u < K; ++u) S[u] += fabsf(a - row[u]);
        }
      }
    }
    if (vi < K) {
      const int dv = dv0 + g;
      const bool rowok = (y + dv >= 0) && (y + dv < H);
      float mv = 9999999.f;
      int mi = 0;
#pragma unroll
      for (int u = 0; u < K; ++u) {
        const int xx = x + ou + u - hs;
        const float sv = (rowok && xx >= 0 && xx < W) ? S[u] : kFill;
        if (sv < mv) { mv = sv; mi = u; }
        if (sv < bmin[u]) { bmin[u] = sv; barg[u] = vi; }
      }
      if (y < H && x < W) {
        const int64_t o = (((int64_t)n * H + y) * W + x) * K + vi;
        cv_v[o] = mv;
        v_star[o] = (float)(mi + bu * sws);
      }
    }
  }
  // merge the groups' u-projections in group 0 (smaller vi wins among equal minima)
  __syncthreads();  // the staging tiles are dead from here on; their space takes the partials
  if (g > 0) {
#pragma unroll
    for (int u = 0; u < K; ++u) {
      m_v[g - 1][pl][u] = bmin[u];
      m_i[g - 1][pl][u] = (unsigned char)barg[u];
    }
  }
  __syncthreads();
  if (g == 0 && y < H && x < W) {
    const int64_t o = (((int64_t)n * H + y) * W + x) * K;
#pragma unroll
    for (int u = 0; u < K; ++u) {
      float bv_ = bmin[u];
      int bi = barg[u];
#pragma unroll
      for (int gg = 1; gg < G; ++gg) {
        const float v2 = m_v[gg - 1][pl][u];
        const int i2 = m_i[gg - 1][pl][u];
        if (v2 < bv_ || (v2 == bv_ && i2 < bi)) { bv_ = v2; bi = i2; }
      }
      cv_u[o + u] = bv_;
      u_star[o + u] = (float)(bi + bv * sws);
    }
  }
}

// Any window size: the block layout and staging of sad_flow_fwd_k with the horizontal
// displacements taken in chunks of FT_UC (S, and the running u-projection, for one chunk in
// registers).  The v-projection of a chunk is merged into cv_v / v_star by the thread that owns
// the (pixel, v) entry: a later chunk replaces it only when strictly smaller, so the first u
// index of the minimum wins as in the reference's scan.  The vertical displacements run over
// [-h, h] (h = sws / 2): for odd sws the row v_idx = K - 1 keeps the fill value.
constexpr int FT_UC = 32;

__global__ __launch_bounds__(512) void sad_flow_fwd_tiled_k(const float *__restrict__ f0, const float *__restrict__ f1,
                                                            float *__restrict__ cv_u, float *__restrict__ cv_v,
                                                            float *__restrict__ u_star, float *__restrict__ v_star, int C,
                                                            int H, int W, int sws, int ou, int ov, int bu, int bv) {
  constexpr int G = FL_G, UC = FT_UC;
  constexpr int SW = FL_TX + UC - 1;  // staged f1 columns
  constexpr int SR = FL_TY + G - 1;   // staged f1 rows (4 consecutive displacements)
  constexpr int NP = FL_TY * FL_TX;
  constexpr size_t kStage = sizeof(float) * (FL_CB * SR * SW + FL_CB * FL_TY * FL_TX);
  constexpr size_t kMerge = (sizeof(float) + sizeof(int)) * (G - 1) * NP * UC;
  __shared__ __align__(16) unsigned char lds[kStage > kMerge ? kStage : kMerge];
  auto s1 = reinterpret_cast<float (*)[SR][SW]>(lds);
  auto s0 = reinterpret_cast<float (*)[FL_TY][FL_TX]>(lds + sizeof(float) * FL_CB * SR * SW);
  auto m_v = reinterpret_cast<float (*)[NP][UC]>(lds);
  auto m_i = reinterpret_cast<int (*)[NP][UC]>(lds + sizeof(float) * (G - 1) * NP * UC);
  const int K = sws + 1, hs = sws / 2;
  const int NI = (K + G - 1) / G;
  const int x0 = blockIdx.x * FL_TX;
  const int y0 = blockIdx.y * FL_TY;
  const int n = blockIdx.z;
  const int t = threadIdx.x;
  const int g = t / NP;
  const int pl = t - g * NP;
  const int ty = pl / FL_TX, tx = pl - ty * FL_TX;
  const int y = y0 + ty, x = x0 + tx;
  const bool inside = y < H && x < W;
  const int64_t HW = (int64_t)H * W;
  const float *F0 = f0 + (int64_t)n * C * HW;
  const float *F1 = f1 + (int64_t)n * C * HW;
  const int64_t obase = (((int64_t)n * H + y) * W + x) * K;

  for (int u0 = 0; u0 < K; u0 += UC) {
    const int nu = min(UC, K - u0);
    float bmin[UC];
    int barg[UC];
#pragma unroll
    for (int u = 0; u < UC; ++u) { bmin[u] = 9999999.f; barg[u] = 0; }
    for (int i = 0; i < NI; ++i) {
      const int vi = g + G * i;
      const int dv0 = ov + G * i - hs;  // displacement of group 0 at this round
      float S[UC];
#pragma unroll
      for (int u = 0; u < UC; ++u) S[u] = 0.f;
      for (int c0 = 0; c0 < C; c0 += FL_CB) {
        __syncthreads();
        for (int e = t; e < FL_CB * SR * SW; e += G * NP) {
          const int cc = e / (SR * SW);
          const int r = e - cc * SR * SW;
          const int ry = r / SW, j = r - ry * SW;
          const int c = c0 + cc, yy = y0 + ry + dv0, xx = x0 + ou - hs + u0 + j;
          s1[cc][ry][j] = (c < C && yy >= 0 && yy < H && xx >= 0 && xx < W) ? F1[(int64_t)c * HW + (int64_t)yy * W + xx] : 0.f;
        }
        for (int e = t; e < FL_CB * NP; e += G * NP) {
          const int cc = e / NP;
          const int r = e - cc * NP;
          const int ry = r / FL_TX, j = r - ry * FL_TX;
          const int c = c0 + cc, yy = y0 + ry, xx = x0 + j;
          s0[cc][ry][j] = (c < C && yy < H && xx < W) ? F0[(int64_t)c * HW + (int64_t)yy * W + xx] : 0.f;
        }
        __syncthreads();
        if (vi < K) {
          const int nc = min(FL_CB, C - c0);
          for (int cc = 0; cc < nc; ++cc) {
            const float a = s0[cc][ty][tx];
            const float *row = &s1[cc][ty + g][tx];
#pragma unroll
            for (int u = 0; u < UC; ++u) S[u] += fabsf(a - row[u]);
          }
        }
      }
      if (vi < K) {
        const int dv = dv0 + g;
        const bool rowok = vi <= 2 * hs && y + dv >= 0 && y + dv < H;
        float mv = 9999999.f;
        int mi = 0;
#pragma unroll
        for (int u = 0; u < UC; ++u) {
          if (u < nu) {
            const int xx = x + ou + u0 + u - hs;
            const float sv = (rowok && xx >= 0 && xx < W) ? S[u] : kFill;
            if (sv < mv) { mv = sv; mi = u0 + u; }
            if (sv < bmin[u]) { bmin[u] = sv; barg[u] = vi; }
          }
        }
        if (inside) {
          const int64_t o = obase + vi;
          if (u0 == 0 || mv < cv_v[o]) {
            cv_v[o] = mv;
            v_star[o] = (float)(mi + bu * sws);
          }
        }
      }
    }
    // merge the groups' u-projections of this chunk in group 0 (smaller vi wins ties)
    __syncthreads();  // the staging tiles are dead from here on; their space takes the partials
    if (g > 0) {
#pragma unroll
      for (int u = 0; u < UC; ++u) {
        m_v[g - 1][pl][u] = bmin[u];
        m_i[g - 1][pl][u] = barg[u];
      }
    }
    __syncthreads();
    if (g == 0 && inside) {
#pragma unroll
      for (int u = 0; u < UC; ++u) {
        if (u < nu) {
          float bv_ = bmin[u];
          int bi = barg[u];
#pragma unroll
          for (int gg = 1; gg < G; ++gg) {
            const float v2 = m_v[gg - 1][pl][u];
            const int i2 = m_i[gg - 1][pl][u];
            if (v2 < bv_ || (v2 == bv_ && i2 < bi)) { bv_ = v2; bi = i2; }
          }
          cv_u[obase + u0 + u] = bv_;
          u_star[obase + u0 + u] = (float)(bi + bv * sws);
        }
      }
    }
  }
}

// ----------------------------------------------------------------------------- flow bwd
// Gradient of the min-projected volumes with the argmins fixed (flow_mp_sad_kernel.cu:126-200).
// Labels k in [0, 2h] carry gradient (h = sws / 2; for odd sws the last label has none, as
// the reference's u/v loops run over [-h, h]).  For label k of pixel p the stored argmin s
// gives the other displacement e = s - h; with (u, v) = (k - h, e) for the u-volume and
// (e, k - h) for the v-volume, a match q = p + (u, v) inside the image contributes
//   df0[p] += sign(f0[p] - f1[q]) * g,   df1[q] -= sign(f0[p] - f1[q]) * g.
// df0 is a gather along the pixel's own argmins.  df1 is a gather too (deterministic, unlike
// the reference's atomics): for stored argmins inside the window (0 <= s < K) the sources
// landing on a pixel lie within K - 1 pixels of it, and are found as below; argmins outside
// the window (only from block offsets) are added afterwards by sad_flow_bwd_far_k.  Terms
// are summed per pixel in a fixed order: u-volume then v-volume, label ascending, e ascending
// (df1's extra landings in pieces whose partial sums are added in that order).

// K <= 32.  Block = 8 x 32 pixels of one image.  Every source a pixel's terms read lies within
// E = K - 1 - h pixels of it (in-window argmins), so the block works on a region of the tile
// plus an E halo:
//   1. df0, per volume: each thread keeps its own terms (region offset of the match and
//      gradient, per label) in registers; channel chunk by chunk, f1 of the region is staged
//      in LDS and the own terms are summed from it;
//   2. df1, per volume: the argmins of the region's sources are read coalesced (lane =
//      label, 8 pixel columns per step) and OR-ed into per-target landing masks in LDS --
//      bit s of mask[k][target] says "the source at target - (u, v) with label k and argmin
//      s lands here"; LDS OR is order-free, so the masks are deterministic.  The first
//      landing of each label (usually the only one) is resolved into registers, the others
//      (displacement, label and gradient) into the block's LDS pool, each thread's extra
//      landings in one contiguous range (start = exclusive prefix sum of the threads' counts, so
//      the placement is deterministic).  Extra landings are uneven: where the argmins of a
//      label are far from the flow they land about Poisson(1) times per target, and where the
//      flow converges many sources share a target (the 256x512, sws 24 workload: 10.7 extra
//      landings per pixel on average, up to 155; up to 5184 in one block), so the pool takes
//      what the LDS budget leaves (5760 entries at sws 24).  Then, channel chunk by chunk with
//      f0 of the region staged in LDS (in the masks' memory), the first landings are summed
//      label by label, then the pooled ones in order (label, then argmin ascending): each
//      thread sums the first L of its own (L = the block's pooled landings / 256, rounded up)
//      and the rest go in pieces of L to helper threads, whose partial sums the owner adds in
//      the pieces' order, so that no thread sums more than 2L however the landings fall; a
//      thread whose landings overflow the pool adds the rest after the last chunk, channel by
//      channel, found again from the argmin volume (they come last in each channel's sum).
// The u-volume's partial sums pass through df0 / df1 (same thread).  Own terms whose argmin
// lies outside the window read f1 from global memory in place.
constexpr int FM_TX = 32, FM_TY = 8, FM_NT = FM_TX * FM_TY, FM_KMAX = 32, FM_CH = 4;
constexpr int FM_NONE = -1;  // own term without a match inside the image
// LDS per block: two blocks (eight waves) share a CU
constexpr int FM_LDS_BUDGET = 78 * 1024;

// bytes of the LDS area shared by the landing masks [KL][FM_NT] and the staged region [CH][RR][RC]
int flow_bwd_union_bytes(int sws, int ch) {
  const int K = sws + 1, hs = sws / 2, KL = 2 * hs + 1, E = K - 1 - hs;
  const int m = 4 * KL * FM_NT, r = 4 * ch * (FM_TY + 2 * E) * (FM_TX + 2 * E);
  return (m > r ? m : r + 15) & ~15;
}
// the pooled landings' pieces: one partial sum per channel of the chunk and one descriptor
// per piece, at most one piece per thread
int flow_bwd_piece_bytes(int ch) { return FM_NT * (4 * ch + 4); }
// entries of the pool of extra landings (a float gradient and a packed offset each)
int flow_bwd_pool(int sws, int ch) {
  const int left = (FM_LDS_BUDGET - flow_bwd_union_bytes(sws, ch) - flow_bwd_piece_bytes(ch)) / 8;
  return left > 4 * FM_NT ? left : 4 * FM_NT;
}
size_t flow_bwd_lds_bytes(int sws, int ch) {
  return (size_t)flow_bwd_union_bytes(sws, ch) + flow_bwd_piece_bytes(ch) + 8 * (size_t)flow_bwd_pool(sws, ch);
}

// region rows of nc channel planes (buffer resource over them) -> reg[CH][RR][RC]; a wave
// takes one row (RC <= 64) per step, 8 steps' loads in flight before their stores; zeros
// outside the image and past nc channels (dropped loads read 0)
__device__ __forceinline__ void stage_region(float *reg, __amdgpu_buffer_rsrc_t F, int nc, int nrows, int RR, int RC,
                                             int ry0, int rx0, int H, int W, int t) {
  constexpr int NW = FM_NT / 64, B = 8;
  const int c = t & 63, w = t >> 6;
  const int xs = rx0 + c;
  const bool cok = c < RC && xs >= 0 && xs < W;
  for (int b0 = 0; b0 < nrows; b0 += NW * B) {
    float v[B];
#pragma unroll
    for (int q = 0; q < B; ++q) {
      const int jr = b0 + NW * q + w;
      const int j = jr / RR, r = jr - j * RR;  // wave-uniform
      const int ys = ry0 + r;
      const bool ok = cok && jr < nrows && j < nc && ys >= 0 && ys < H;
      v[q] = buf_ld(F, ok ? (uint32_t)(((j * H + ys) * W + xs) * 4) : kBufDrop, 0);
    }
#pragma unroll
    for (int q = 0; q < B; ++q) {
      const int jr = b0 + NW * q + w;
      if (c < RC && jr < nrows) reg[jr * RC + c] = v[q];
    }
  }
}

// KM >= K: the size of the per-label register arrays (16, 25 or 32: fewer registers for
// smaller windows)
template <int CH, int KM>
__global__ __launch_bounds__(FM_NT) void sad_flow_bwd_lds_k(const float *__restrict__ f0, const float *__restrict__ f1,
                                                             const float *__restrict__ gu, const float *__restrict__ gv,
                                                             const float *__restrict__ us, const float *__restrict__ vs,
                                                             float *__restrict__ df0, float *__restrict__ df1, int C,
                                                             int H, int W, int sws, int ubytes, int pool) {
  extern __shared__ __align__(16) uint32_t fm_lds[];
  const int K = sws + 1, hs = sws / 2, KL = 2 * hs + 1, E = K - 1 - hs;
  const int RR = FM_TY + 2 * E, RC = FM_TX + 2 * E, RN = RR * RC;
  // the masks and the staged region share the first ubytes (never live together)
  uint32_t *mask = fm_lds;                                                 // [KL][FM_NT]
  float *reg = reinterpret_cast<float *>(fm_lds);                          // [CH][RR][RC]
  float *psum = reinterpret_cast<float *>(fm_lds + ubytes / 4);            // piece partials [FM_NT][CH]
  uint32_t *desc = reinterpret_cast<uint32_t *>(psum + FM_NT * CH);        // piece descriptors [FM_NT]
  float *xg = reinterpret_cast<float *>(desc + FM_NT);                     // pool gradients [pool]
  // pool offsets: v+32 | (u+32) << 6 | k << 12, then the source's region offset
  uint32_t *xo = reinterpret_cast<uint32_t *>(xg + pool);
  const int t = threadIdx.x;
  const int ty = t / FM_TX, tx = t - ty * FM_TX;
  // tiles go to the XCDs in contiguous ranges (bands of tile rows), so that the halo rows
  // that neighbouring tiles both stage are read through one L2
  const int gx = (int)gridDim.x;
  const int tb = xcd_contiguous_block((int)blockIdx.x + gx * (int)blockIdx.y, gx * (int)gridDim.y);
  const int x0 = (tb % gx) * FM_TX, y0 = (tb / gx) * FM_TY, n = blockIdx.z;
  const int y = y0 + ty, x = x0 + tx;
  const int ry0 = y0 - E, rx0 = x0 - E;  // region origin
  const bool inside = y < H && x < W;
  const int64_t HW = (int64_t)H * W;
  const int64_t img = (int64_t)n * HW;
  const int64_t pix = inside ? (int64_t)y * W + x : 0;
  const int own = (ty + E) * RC + (tx + E);  // own pixel in the region
  const int ipix = (int)pix;                 // own pixel in the image
  // the argmin and gradient volumes of this image as buffers: 32-bit offsets, dropped
  // (zero-reading) lanes instead of branches around loads
  const uint32_t vb = (uint32_t)(HW * K * 4);

  // 1. df0: own terms (region offset >= 0, FM_NONE, or -2 - image offset for a match outside
  // the window)
  for (int part = 0; part < 2; ++part) {
    const __amdgpu_buffer_rsrc_t rst = make_rsrc((part ? vs : us) + img * K, vb);
    const __amdgpu_buffer_rsrc_t rgr = make_rsrc((part ? gv : gu) + img * K, vb);
    int toff[KM];
    float tg[KM];
    bool far = false;
#pragma unroll
    for (int k = 0; k < KM; ++k) {
      const uint32_t o = inside && k < KL ? (uint32_t)((ipix * K + k) * 4) : kBufDrop;
      const int sv = (int)buf_ld(rst, o, 0);
      const float g = buf_ld(rgr, o, 0);
      const int m = sv - hs;
      const int u = part ? m : k - hs, v = part ? k - hs : m;
      const int xx = x + u, yy = y + v;
      const bool ok = inside && k < KL && xx >= 0 && xx < W && yy >= 0 && yy < H;
      const bool win = (unsigned)sv < (unsigned)K;
      toff[k] = !ok ? FM_NONE : win ? own + v * RC + u : -2 - (yy * W + xx);
      tg[k] = ok ? g : 0.f;
      far |= ok && !win;
    }
    for (int c0 = 0; c0 < C; c0 += CH) {
      const int nc = min(CH, C - c0);
      __syncthreads();  // the previous chunk's region is read
      stage_region(reg, make_rsrc(f1 + ((int64_t)n * C + c0) * HW, (uint32_t)(nc * HW * 4)), nc, CH * RR, RR, RC, ry0,
                   rx0, H, W, t);
      __syncthreads();
      if (inside) {
        float a[CH], r0[CH];
#pragma unroll
        for (int j = 0; j < CH; ++j) {
          const int64_t o = ((int64_t)n * C + c0 + min(j, nc - 1)) * HW + pix;
          a[j] = f0[o];
          r0[j] = part ? df0[o] : 0.f;
        }
#pragma unroll
        for (int k = 0; k < KM; ++k) {
          const int off = toff[k];
          const float g = tg[k];
          if (off >= 0) {
#pragma unroll
            for (int j = 0; j < CH; ++j) r0[j] += sgn_eps(a[j] - reg[j * RN + off]) * g;
          } else if (far && off != FM_NONE) {
            const int64_t q = -2 - (int64_t)off;
#pragma unroll
            for (int j = 0; j < CH; ++j)
              if (j < nc) r0[j] += sgn_eps(a[j] - f1[((int64_t)n * C + c0 + j) * HW + q]) * g;
          }
        }
#pragma unroll
        for (int j = 0; j < CH; ++j)
          if (j < nc) df0[((int64_t)n * C + c0 + j) * HW + pix] = r0[j];
      }
    }
  }

  // 2. df1
  const int lk = t & 31, cgp = t >> 5;  // scatter: lane label, pixel column group
  for (int part = 0; part < 2; ++part) {
    const float *gg = part ? gv : gu;
    const __amdgpu_buffer_rsrc_t rst = make_rsrc((part ? vs : us) + img * K, vb);
    const __amdgpu_buffer_rsrc_t rgr = make_rsrc(gg + img * K, vb);
    __syncthreads();  // the previous masks and region are read
    for (int k = 0; k < KL; ++k) mask[k * FM_NT + t] = 0u;
    __syncthreads();
    if (lk < KL) {
      for (int r = 0; r < RR; r += 2) {
        int sr[2][8];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int ys = ry0 + r + h;
          const bool rok = r + h < RR && ys >= 0 && ys < H;
#pragma unroll
          for (int ci = 0; ci < 8; ++ci) {
            const int c = cgp + 8 * ci, xs = rx0 + c;
            const bool ok = rok && c < RC && xs >= 0 && xs < W;
            const int sv = (int)buf_ld(rst, ok ? (uint32_t)(((ys * W + xs) * K + lk) * 4) : kBufDrop, 0);
            sr[h][ci] = ok ? sv : -1;
          }
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
#pragma unroll
          for (int ci = 0; ci < 8; ++ci) {
            const int sv = sr[h][ci];
            if ((unsigned)sv < (unsigned)K) {
              const int ys = ry0 + r + h, xs = rx0 + cgp + 8 * ci;
              const int e = sv - hs;
              const int yt = ys + (part ? lk - hs : e) - y0, xt = xs + (part ? e : lk - hs) - x0;
              if ((unsigned)yt < (unsigned)FM_TY && (unsigned)xt < (unsigned)FM_TX)
                atomicOr(&mask[lk * FM_NT + yt * FM_TX + xt], 1u << sv);
            }
          }
        }
      }
    }
    __syncthreads();
    // first landing per label into registers; count the others
    int foff[KM];
    float fg[KM];
    int nx = 0;  // landings past the first of their label
#pragma unroll
    for (int k = 0; k < KM; ++k) {
      const uint32_t m = inside && k < KL ? mask[k * FM_NT + t] : 0u;
      const int e = (m != 0u ? __builtin_ctz(m) : hs) - hs;
      const int u = part ? e : k - hs, v = part ? k - hs : e;
      const float g = buf_ld(rgr, m != 0u ? (uint32_t)(((ipix - v * W - u) * K + k) * 4) : kBufDrop, 0);
      foff[k] = m != 0u ? own - v * RC - u : FM_NONE;
      fg[k] = m != 0u ? g : 0.f;
      nx += __builtin_popcount(m & (m - 1u));
    }
    // the thread's range of the pool: exclusive prefix sum of the counts over the block (in the
    // pool's gradient area, written only after every thread has read its sum)
    int *scan = reinterpret_cast<int *>(xg);
    scan[t] = nx;
    __syncthreads();
    for (int d = 1; d < FM_NT; d <<= 1) {
      const int add = t >= d ? scan[t - d] : 0;
      __syncthreads();
      scan[t] += add;
      __syncthreads();
    }
    const int x0p = scan[t] - nx;                     // first pool entry of the thread
    const int nfit = min(nx, max(0, pool - x0p));     // of its landings in the pool
    const int npool = min(scan[FM_NT - 1], pool);     // pooled landings of the block
    __syncthreads();
    // the extra landings into the pool in order (label, then argmin ascending)
    {
      int i = 0;
      for (int k = 0; k < KL && i < nfit; ++k) {
        uint32_t m = inside ? mask[k * FM_NT + t] : 0u;
        m &= m - 1u;
        while (m != 0u && i < nfit) {
          const int e2 = __builtin_ctz(m) - hs;
          m &= m - 1u;
          const int u2 = part ? e2 : k - hs, v2 = part ? k - hs : e2;
          xo[x0p + i] = (uint32_t)(v2 + 32) | (uint32_t)(u2 + 32) << 6 | (uint32_t)k << 12;
          ++i;
        }
      }
    }
    // the pooled landings' gradients, four loads in flight at a time; each entry's packed
    // displacement is then replaced by the source's region offset (the thread's own slots)
    for (int i0 = 0; i0 < nfit; i0 += 4) {
      float g4[4];
      int o4[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const bool ok = i0 + q < nfit;
        const uint32_t w = ok ? xo[x0p + i0 + q] : 0u;
        const int v = (int)(w & 63u) - 32, u = (int)((w >> 6) & 63u) - 32, k = (int)(w >> 12);
        g4[q] = buf_ld(rgr, ok ? (uint32_t)(((ipix - v * W - u) * K + k) * 4) : kBufDrop, 0);
        o4[q] = own - v * RC - u;
      }
#pragma unroll
      for (int q = 0; q < 4; ++q)
        if (i0 + q < nfit) {
          xg[x0p + i0 + q] = g4[q];
          xo[x0p + i0 + q] = (uint32_t)o4[q];
        }
    }
    // Pieces: a thread sums the first L of its pooled landings itself and hands the rest, in
    // pieces of L, to the threads numbered by the pieces' order; with L = ceil(npool / FM_NT)
    // the pieces number fewer than npool / L <= FM_NT (one per thread at most), and no thread
    // sums more than 2L pooled landings per chunk, however unevenly they fall on the pixels.
    // Piece descriptors: first pool entry | length << 16 | owner thread << 24.
    const int L = max(1, (npool + FM_NT - 1) / FM_NT);
    const int nown = min(nfit, L);
    const int npc = (nfit - nown + L - 1) / L;
    int *pscan = reinterpret_cast<int *>(psum);
    pscan[t] = npc;
    __syncthreads();
    for (int d = 1; d < FM_NT; d <<= 1) {
      const int add = t >= d ? pscan[t - d] : 0;
      __syncthreads();
      pscan[t] += add;
      __syncthreads();
    }
    const int ps = pscan[t] - npc;         // the thread's first piece
    const int npieces = pscan[FM_NT - 1];  // the block's pieces
    for (int i = 0; i < npc; ++i)
      desc[ps + i] = (uint32_t)(x0p + nown + i * L) | (uint32_t)min(L, nfit - nown - i * L) << 16 | (uint32_t)t << 24;
    for (int c0 = 0; c0 < C; c0 += CH) {
      const int nc = min(CH, C - c0);
      __syncthreads();  // the previous chunk's region and partials are read; descriptors written
      stage_region(reg, make_rsrc(f0 + ((int64_t)n * C + c0) * HW, (uint32_t)(nc * HW * 4)), nc, CH * RR, RR, RC, ry0,
                   rx0, H, W, t);
      __syncthreads();
      float b[CH] = {}, r1[CH] = {};
      if (inside) {
#pragma unroll
        for (int j = 0; j < CH; ++j) {
          const int64_t o = ((int64_t)n * C + c0 + min(j, nc - 1)) * HW + pix;
          b[j] = f1[o];
          r1[j] = part ? df1[o] : 0.f;
        }
#pragma unroll
        for (int k = 0; k < KM; ++k) {
          const int off = foff[k];
          if (off >= 0) {
            const float g = fg[k];
#pragma unroll
            for (int j = 0; j < CH; ++j) r1[j] -= sgn_eps(reg[j * RN + off] - b[j]) * g;
          }
        }
        for (int i = 0; i < nown; ++i) {
          const int off = (int)xo[x0p + i];
          const float g = xg[x0p + i];
#pragma unroll
          for (int j = 0; j < CH; ++j) r1[j] -= sgn_eps(reg[j * RN + off] - b[j]) * g;
        }
      }
      if (t < npieces) {  // the piece numbered t, for its owner's pixel
        const uint32_t dd = desc[t];
        const int s0 = (int)(dd & 0xFFFFu), len = (int)((dd >> 16) & 0xFFu), o = (int)(dd >> 24);
        const int64_t po = (int64_t)(y0 + o / FM_TX) * W + (x0 + o % FM_TX);
        float bo[CH], pp[CH];
#pragma unroll
        for (int j = 0; j < CH; ++j) {
          bo[j] = f1[((int64_t)n * C + c0 + min(j, nc - 1)) * HW + po];
          pp[j] = 0.f;
        }
        for (int i = 0; i < len; ++i) {
          const int off = (int)xo[s0 + i];
          const float g = xg[s0 + i];
#pragma unroll
          for (int j = 0; j < CH; ++j) pp[j] += sgn_eps(reg[j * RN + off] - bo[j]) * g;
        }
#pragma unroll
        for (int j = 0; j < CH; ++j) psum[t * CH + j] = pp[j];
      }
      __syncthreads();  // the partials are written
      if (inside) {
        for (int q = ps; q < ps + npc; ++q) {
#pragma unroll
          for (int j = 0; j < CH; ++j) r1[j] -= psum[q * CH + j];
        }
#pragma unroll
        for (int j = 0; j < CH; ++j)
          if (j < nc) df1[((int64_t)n * C + c0 + j) * HW + pix] = r1[j];
      }
    }
    // rare (the block's pool is full): the landings past it, in the same order -- label, then
    // argmin ascending, the first of each label and the pooled ones skipped -- after every chunk
    if (inside && nx > nfit) {
      const float *st = (part ? vs : us) + img * K;
      int ix = 0;
      for (int k = 0; k < KL; ++k) {
        bool first = true;
        for (int sv = 0; sv < K; ++sv) {
          const int e = sv - hs;
          const int u = part ? e : k - hs, v = part ? k - hs : e;
          const int ys = y - v, xs = x - u;
          if (ys < 0 || ys >= H || xs < 0 || xs >= W || (int)st[((int64_t)ys * W + xs) * K + k] != sv) continue;
          if (first) {
            first = false;
            continue;
          }
          if (ix++ < nfit) continue;
          const float g = gg[(img + (int64_t)ys * W + xs) * K + k];
          for (int c = 0; c < C; ++c) {
            const int64_t o = ((int64_t)n * C + c) * HW;
            df1[o + pix] -= sgn_eps(f0[o + (int64_t)ys * W + xs] - f1[o + pix]) * g;
          }
        }
      }
    }
  }
}

// K > 32: one thread per pixel and chunk of FB_CB channels, the sources landing on the pixel
// found by a K x K search of the argmins.  The sources a lane finds are first collected in
// search order in a per-thread LDS list; the channel updates then run over the lists in
// lockstep.  A list that overflows FB_CAP is finished by a second pass over the remaining
// landings.
constexpr int FB_CB = 32;
constexpr int FB_CAP = 48;
constexpr int FB_THREADS = 128;

__global__ __launch_bounds__(FB_THREADS) void sad_flow_bwd_search_k(const float *__restrict__ f0, const float *__restrict__ f1,
                                                                    const float *__restrict__ gu, const float *__restrict__ gv,
                                                                    const float *__restrict__ us, const float *__restrict__ vs,
                                                                    float *__restrict__ df0, float *__restrict__ df1,
                                                                    int N, int C, int H, int W, int sws) {
  __shared__ int l_src[FB_CAP][FB_THREADS];
  __shared__ float l_g[FB_CAP][FB_THREADS];
  const int K = sws + 1, hs = sws / 2, KL = 2 * hs + 1, E = K - 1 - hs;
  const int64_t HW = (int64_t)H * W;
  const int64_t total = (int64_t)N * HW;
  const int t = threadIdx.x;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + t; p < total; p += (int64_t)gridDim.x * blockDim.x) {
    const int n = (int)(p / HW);
    const int y = (int)((p / W) % H), x = (int)(p % W);
    const int64_t pix = p - (int64_t)n * HW;
    const int64_t img = (int64_t)n * HW;
    for (int c0 = 0; c0 < C; c0 += FB_CB) {
      const int nc = min(FB_CB, C - c0);
      const float *F0 = f0 + ((int64_t)n * C + c0) * HW;
      const float *F1 = f1 + ((int64_t)n * C + c0) * HW;
      float a[FB_CB], b[FB_CB], r0[FB_CB], r1[FB_CB];
#pragma unroll
      for (int j = 0; j < FB_CB; ++j) {
        const bool ok = j < nc;
        a[j] = ok ? F0[(int64_t)j * HW + pix] : 0.f;
        b[j] = ok ? F1[(int64_t)j * HW + pix] : 0.f;
        r0[j] = 0.f;
        r1[j] = 0.f;
      }
      for (int part = 0; part < 2; ++part) {
        const float *st = part ? vs : us;
        const float *gg = part ? gv : gu;
        for (int k = 0; k < KL; ++k) {
          const int64_t o = p * K + k;
          const int m = (int)st[o] - hs;
          const int u = part ? m : k - hs, v = part ? k - hs : m;
          const int xx = x + u, yy = y + v;
          if (xx < 0 || xx >= W || yy < 0 || yy >= H) continue;
          const float g = gg[o];
          const int64_t src = (int64_t)yy * W + xx;
#pragma unroll
          for (int j = 0; j < FB_CB; ++j)
            if (j < nc) r0[j] += sgn_eps(a[j] - F1[(int64_t)j * HW + src]) * g;
        }
      }
      for (int part = 0; part < 2; ++part) {
        const float *st = part ? vs : us;
        const float *gg = part ? gv : gu;
        int cnt = 0;
        for (int k = 0; k < KL; ++k) {
          for (int e = -hs; e <= E; ++e) {
            // u-volume: label k is the horizontal displacement, e the vertical one; v: swapped
            const int u = part ? e : k - hs, v = part ? k - hs : e;
            const int xs = x - u, ys = y - v;
            if (xs < 0 || xs >= W || ys < 0 || ys >= H) continue;
            const int64_t src = (int64_t)ys * W + xs;
            const int64_t o = (img + src) * K + k;
            if ((int)st[o] - hs != e) continue;
            if (cnt < FB_CAP) {
              l_src[cnt][t] = (int)src;
              l_g[cnt][t] = gg[o];
            }
            ++cnt;
          }
        }
        const int m = min(cnt, FB_CAP);
        for (int i = 0; i < m; ++i) {
          const int64_t src = l_src[i][t];
          const float g = l_g[i][t];
#pragma unroll
          for (int j = 0; j < FB_CB; ++j)
            if (j < nc) r1[j] -= sgn_eps(F0[(int64_t)j * HW + src] - b[j]) * g;
        }
        if (cnt > FB_CAP) {
          int seen = 0;
          for (int k = 0; k < KL; ++k) {
            for (int e = -hs; e <= E; ++e) {
              const int u = part ? e : k - hs, v = part ? k - hs : e;
              const int xs = x - u, ys = y - v;
              if (xs < 0 || xs >= W || ys < 0 || ys >= H) continue;
              const int64_t src = (int64_t)ys * W + xs;
              const int64_t o = (img + src) * K + k;
              if ((int)st[o] - hs != e) continue;
              if (seen++ < FB_CAP) continue;
              const float g = gg[o];
#pragma unroll
              for (int j = 0; j < FB_CB; ++j)
                if (j < nc) r1[j] -= sgn_eps(F0[(int64_t)j * HW + src] - b[j]) * g;
            }
          }
        }
      }
      float *D0 = df0 + ((int64_t)n * C + c0) * HW;
      float *D1 = df1 + ((int64_t)n * C + c0) * HW;
#pragma unroll
      for (int j = 0; j < FB_CB; ++j)
        if (j < nc) {
          D0[(int64_t)j * HW + pix] = r0[j];
          D1[(int64_t)j * HW + pix] = r1[j];
        }
    }
  }
}

// Argmins outside the window (s < 0 or s >= K; only block-offset argmins have them): their
// df1 terms, added after the gather kernels with atomics (the reference's own scheme).  One
// thread per pixel reads the pixel's argmins (in-window ones are skipped).
__global__ __launch_bounds__(256) void sad_flow_bwd_far_k(const float *__restrict__ f0, const float *__restrict__ f1,
                                                          const float *__restrict__ gu, const float *__restrict__ gv,
                                                          const float *__restrict__ us, const float *__restrict__ vs,
                                                          float *__restrict__ df1, int N, int C, int H, int W, int sws) {
  const int K = sws + 1, hs = sws / 2, KL = 2 * hs + 1;
  const int64_t HW = (int64_t)H * W;
  const int64_t total = (int64_t)N * HW;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < total; p += (int64_t)gridDim.x * blockDim.x) {
    for (int part = 0; part < 2; ++part) {
      const float *st = part ? vs : us;
      for (int k = 0; k < KL; ++k) {
        const int sv = (int)st[p * K + k];
        if ((unsigned)sv < (unsigned)K) continue;
        const int n = (int)(p / HW);
        const int64_t pix = p - (int64_t)n * HW;
        const int y = (int)(pix / W), x = (int)(pix - (int64_t)y * W);
        const int e = sv - hs;
        const int u = part ? e : k - hs, v = part ? k - hs : e;
        const int xx = x + u, yy = y + v;
        if (xx < 0 || xx >= W || yy < 0 || yy >= H) continue;
        const float g = (part ? gv : gu)[p * K + k];
        const int64_t q = (int64_t)yy * W + xx;
        for (int c = 0; c < C; ++c) {
          const int64_t plane = ((int64_t)n * C + c) * HW;
          atomicAdd(&df1[plane + q], -(sgn_eps(f0[plane + pix] - f1[plane + q]) * g));
        }
      }
    }
  }
}

unsigned grid_for(int64_t total, int threads) {
  int64_t b = (total + threads - 1) / threads;
  if (b > 1 << 20) b = 1 << 20;
  if (b < 1) b = 1;
  return (unsigned)b;
}

}  // namespace
}  // namespace bpl

using namespace bpl;

extern "C" int bpl_sad_stereo_fwd(const float *f0, const float *f1, float *cost, int64_t N, int64_t C, int64_t H,
                                  int64_t W, int min_disp, int max_disp, float step, void *stream) {
  BPL_REQUIRE(f0 && f1 && cost, BPL_E_ARG, "bpl_sad_stereo_fwd: null pointer");
  BPL_REQUIRE(N >= 1 && C >= 1 && H >= 1 && W >= 1, BPL_E_ARG, "bpl_sad_stereo_fwd: sizes must be positive");
  BPL_REQUIRE(step > 0.f, BPL_E_ARG, "bpl_sad_stereo_fwd: step must be positive");
  BPL_REQUIRE(H < 65536 && N * C * H * W < (1LL << 40), BPL_E_LIMIT, "bpl_sad_stereo_fwd: tensor too large");
  const int D = (int)((float)(max_disp - min_disp + 1) / step);
  BPL_REQUIRE(D >= 1, BPL_E_ARG, "bpl_sad_stereo_fwd: empty disparity range");
  hipStream_t st = as_stream(stream);
  KernelTimer kt(BPL_KERNEL_SAD_STEREO_FWD, st);
  if (step == 1.0f) {
    const int dpasses = (D + ST_TD - 1) / ST_TD;
    BPL_REQUIRE(N * dpasses < 65536, BPL_E_LIMIT, "bpl_sad_stereo_fwd: N * ceil(D/64) too large");
    dim3 grid((unsigned)((W + ST_TX - 1) / ST_TX), (unsigned)H, (unsigned)(N * dpasses));
    hipLaunchKernelGGL(sad_stereo_fwd_k, grid, dim3(256), 0, st, f0, f1, cost, (int)C, (int)H, (int)W, D, min_disp,
                       dpasses);
  } else {
    hipLaunchKernelGGL(sad_stereo_fwd_step_k, dim3(grid_for(N * H * W * D, 256)), dim3(256), 0, st, f0, f1, cost,
                       (int)N, (int)C, (int)H, (int)W, D, min_disp, step);
  }
  return check_hip(hipGetLastError(), "bpl_sad_stereo_fwd launch");
}

extern "C" int bpl_sad_stereo_softmax_fwd(const float *f0, const float *f1, float *y, int64_t N, int64_t C, int64_t H,
                                          int64_t W, int min_disp, int max_disp, float scale, float sign, void *stream) {
  BPL_REQUIRE(f0 && f1 && y, BPL_E_ARG, "bpl_sad_stereo_softmax_fwd: null pointer");
  BPL_REQUIRE(N >= 1 && C >= 1 && H >= 1 && W >= 1, BPL_E_ARG, "bpl_sad_stereo_softmax_fwd: sizes must be positive");
  const int D = max_disp - min_disp + 1;
  BPL_REQUIRE(D >= 1, BPL_E_ARG, "bpl_sad_stereo_softmax_fwd: empty disparity range");
  BPL_REQUIRE(D <= 256, BPL_E_LIMIT, "bpl_sad_stereo_softmax_fwd: D=%d exceeds 256", D);
  BPL_REQUIRE(H < 65536 && N < 65536 && N * C * H * W < (1LL << 40), BPL_E_LIMIT,
              "bpl_sad_stereo_softmax_fwd: tensor too large");
  hipStream_t st = as_stream(stream);
  const dim3 grid((unsigned)((W + ST_TX - 1) / ST_TX), (unsigned)H, (unsigned)N);
  const int lpl = (D + 63) / 64;
  KernelTimer kt(BPL_KERNEL_SAD_STEREO_SOFTMAX, st);
  switch (lpl) {
    case 1: hipLaunchKernelGGL(sad_stereo_softmax_k<1>, grid, dim3(256), 0, st, f0, f1, y, (int)C, (int)H, (int)W, D, min_disp, scale, sign); break;
    case 2: hipLaunchKernelGGL(sad_stereo_softmax_k<2>, grid, dim3(256), 0, st, f0, f1, y, (int)C, (int)H, (int)W, D, min_disp, scale, sign); break;
    case 3: hipLaunchKernelGGL(sad_stereo_softmax_k<3>, grid, dim3(256), 0, st, f0, f1, y, (int)C, (int)H, (int)W, D, min_disp, scale, sign); break;
    default: hipLaunchKernelGGL(sad_stereo_softmax_k<4>, grid, dim3(256), 0, st, f0, f1, y, (int)C, (int)H, (int)W, D, min_disp, scale, sign); break;
  }
  return check_hip(hipGetLastError(), "bpl_sad_stereo_softmax_fwd launch");
}

extern "C" int bpl_sad_stereo_bwd(const float *f0, const float *f1, const float *grad_cost, float *df0, float *df1,
                                  int64_t N, int64_t C, int64_t H, int64_t W, int min_disp, int max_disp,
                                  void *stream) {
  BPL_REQUIRE(f0 && f1 && grad_cost && df0 && df1, BPL_E_ARG, "bpl_sad_stereo_bwd: null pointer");
  BPL_REQUIRE(N >= 1 && C >= 1 && H >= 1 && W >= 1, BPL_E_ARG, "bpl_sad_stereo_bwd: sizes must be positive");
  const int D = max_disp - min_disp + 1;
  BPL_REQUIRE(D >= 1, BPL_E_ARG, "bpl_sad_stereo_bwd: empty disparity range");
  BPL_REQUIRE(H < 65536 && N < 65536, BPL_E_LIMIT, "bpl_sad_stereo_bwd: tensor too large");
  dim3 grid((unsigned)((W + SB_TX - 1) / SB_TX), (unsigned)H, (unsigned)N);
  KernelTimer kt(BPL_KERNEL_SAD_STEREO_BWD, as_stream(stream));
  hipLaunchKernelGGL(sad_stereo_bwd_k, grid, dim3(256), 0, as_stream(stream), f0, f1, grad_cost, df0, df1,
                     (int)C, (int)H, (int)W, D, min_disp);
  return check_hip(hipGetLastError(), "bpl_sad_stereo_bwd launch");
}

extern "C" int bpl_sad_flow_fwd(const float *f0, const float *f1, float *cv_u, float *cv_v, float *u_star,
                                float *v_star, int64_t N, int64_t C, int64_t H, int64_t W, int sws, int offset_u,
                                int offset_v, int block_u, int block_v, void *stream) {
  BPL_REQUIRE(f0 && f1 && cv_u && cv_v && u_star && v_star, BPL_E_ARG, "bpl_sad_flow_fwd: null pointer");
  BPL_REQUIRE(N >= 1 && C >= 1 && H >= 1 && W >= 1, BPL_E_ARG, "bpl_sad_flow_fwd: sizes must be positive");
  BPL_REQUIRE(sws >= 0 && sws <= 1024, BPL_E_ARG, "bpl_sad_flow_fwd: sws=%d out of range", sws);
  hipStream_t st = as_stream(stream);
  KernelTimer kt(BPL_KERNEL_SAD_FLOW_FWD, st);
  if (sws == 24) {
    dim3 grid((unsigned)((W + FL_TX - 1) / FL_TX), (unsigned)((H + FL_TY - 1) / FL_TY), (unsigned)N);
    hipLaunchKernelGGL(sad_flow_fwd_k<25>, grid, dim3(FL_G * FL_TY * FL_TX), 0, st, f0, f1, cv_u, cv_v, u_star, v_star, (int)C,
                       (int)H, (int)W, sws, offset_u, offset_v, block_u, block_v);
  } else {
    BPL_REQUIRE(N < 65536 && (H + FL_TY - 1) / FL_TY < 65536, BPL_E_LIMIT, "bpl_sad_flow_fwd: N or H too large");
    dim3 grid((unsigned)((W + FL_TX - 1) / FL_TX), (unsigned)((H + FL_TY - 1) / FL_TY), (unsigned)N);
    hipLaunchKernelGGL(sad_flow_fwd_tiled_k, grid, dim3(FL_G * FL_TY * FL_TX), 0, st, f0, f1, cv_u, cv_v, u_star,
                       v_star, (int)C, (int)H, (int)W, sws, offset_u, offset_v, block_u, block_v);
  }
  return check_hip(hipGetLastError(), "bpl_sad_flow_fwd launch");
}

extern "C" int bpl_sad_flow_bwd(const float *f0, const float *f1, const float *grad_u, const float *grad_v,
                                const float *u_star, const float *v_star, float *df0, float *df1, int64_t N,
                                int64_t C, int64_t H, int64_t W, int sws, void *stream) {
  BPL_REQUIRE(f0 && f1 && grad_u && grad_v && u_star && v_star && df0 && df1, BPL_E_ARG,
              "bpl_sad_flow_bwd: null pointer");
  BPL_REQUIRE(N >= 1 && C >= 1 && H >= 1 && W >= 1, BPL_E_ARG, "bpl_sad_flow_bwd: sizes must be positive");
  BPL_REQUIRE(sws >= 0 && sws <= 1024, BPL_E_ARG, "bpl_sad_flow_bwd: sws=%d out of range", sws);
  hipStream_t st = as_stream(stream);
  KernelTimer kt(BPL_KERNEL_SAD_FLOW_BWD, st);
  // the landing-mask kernel: K <= 32, 32-bit buffer offsets into one image's volumes and
  // feature planes, grid limits; anything else takes the per-pixel search kernel
  const bool masks = sws + 1 <= FM_KMAX && N < 65536 && (H + FM_TY - 1) / FM_TY < 65536 &&
                     H * W * (sws + 1 > FM_CH ? sws + 1 : FM_CH) * 4 < (1LL << 31);
  if (masks) {
    const size_t shm = flow_bwd_lds_bytes(sws, FM_CH);
    const int ubytes = flow_bwd_union_bytes(sws, FM_CH), pool = flow_bwd_pool(sws, FM_CH);
    const dim3 grid((unsigned)((W + FM_TX - 1) / FM_TX), (unsigned)((H + FM_TY - 1) / FM_TY), (unsigned)N);
    auto go = [&](auto km) -> int {
      constexpr int KM = decltype(km)::value;
      if (shm > 64 * 1024) {
        const int rc = check_hip(hipFuncSetAttribute((const void *)sad_flow_bwd_lds_k<FM_CH, KM>,
                                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm),
                                 "bpl_sad_flow_bwd: raising the LDS limit");
        if (rc != BPL_OK) return rc;
      }
      hipLaunchKernelGGL((sad_flow_bwd_lds_k<FM_CH, KM>), grid, dim3(FM_NT), shm, st, f0, f1, grad_u, grad_v, u_star,
                         v_star, df0, df1, (int)C, (int)H, (int)W, sws, ubytes, pool);
      return BPL_OK;
    };
    const int rc = sws + 1 <= 16   ? go(std::integral_constant<int, 16>())
                   : sws + 1 <= 25 ? go(std::integral_constant<int, 25>())
                                   : go(std::integral_constant<int, 32>());
    if (rc != BPL_OK) return rc;
  } else {
    hipLaunchKernelGGL(sad_flow_bwd_search_k, dim3(grid_for(N * H * W, FB_THREADS)), dim3(FB_THREADS), 0, st, f0, f1,
                       grad_u, grad_v, u_star, v_star, df0, df1, (int)N, (int)C, (int)H, (int)W, sws);
  }
  const int rc = check_hip(hipGetLastError(), "bpl_sad_flow_bwd launch");
  if (rc != BPL_OK) return rc;
  hipLaunchKernelGGL(sad_flow_bwd_far_k, dim3(grid_for(N * H * W, 256)), dim3(256), 0, st, f0,
                     f1, grad_u, grad_v, u_star, v_star, df1, (int)N, (int)C, (int)H, (int)W, sws);
  return check_hip(hipGetLastError(), "bpl_sad_flow_bwd far launch");
}
