// Per-pixel label epilogues around the LBP op on gfx950 (one wave per pixel, labels on
// lanes, fixed reduction order: results do not depend on the batch a pixel is solved in).
//
//   bpl_softmax_labels: y = sign * softmax(scale * x) over the contiguous label dimension --
//     the matching softmax softmax(-SAD/T) (reference matching.py:22-25, corenet.py:97-104)
//     followed by the message-passing input cost = -prob (message_passing_op_cuda.py:162);
//   bpl_lbp_beliefs: beliefs = softmin((cost + sum_s M_s) / T) (message_passing_op_cuda.py:170-173)
//     and the winner-take-all label argmax_c beliefs (inference_op.py:20-24, first index).
#include "bpl_common.h"

namespace bpl {
namespace {

constexpr int kMaxLPL = 4;

template <int LPL>
__global__ __launch_bounds__(256) void softmax_labels_k(const float *__restrict__ x, float *__restrict__ y,
                                                        int64_t rows, int C, float scale, float sign) {
  const int lane = threadIdx.x & 63;
  const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= rows) return;  // whole wave exits together
  const float *xr = x + row * C;
  float v[LPL];
  float mx = -__builtin_inff();
#pragma unroll
  for (int k = 0; k < LPL; ++k) {
    const int c = lane + 64 * k;
    v[k] = c < C ? xr[c] * scale : -__builtin_inff();
    mx = fmaxf(mx, v[k]);
  }
  mx = seg_reduce<64>(mx, OpMax());
  float s = 0.f;
#pragma unroll
  for (int k = 0; k < LPL; ++k) {
    const int c = lane + 64 * k;
    v[k] = c < C ? fast_exp(v[k] - mx) : 0.f;
    s += v[k];
  }
  s = seg_reduce<64>(s, OpSum());
  const float r = sign / s;
  float *yr = y + row * C;
#pragma unroll
  for (int k = 0; k < LPL; ++k) {
    const int c = lane + 64 * k;
    if (c < C) yr[c] = v[k] * r;
  }
}

template <int LPL>
__global__ __launch_bounds__(256) void beliefs_k(const float *__restrict__ cost, const float *__restrict__ msg,
                                                 float *__restrict__ beliefs, float *__restrict__ label,
                                                 int N, int H, int W, int C, float inv_T) {
  const int lane = threadIdx.x & 63;
  const int64_t HW = (int64_t)H * W;
  const int64_t p = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (p >= (int64_t)N * HW) return;
  const int64_t n = p / HW, r = p - n * HW;
  const int64_t plane = HW * C;
  const float *m0 = msg + n * 4 * plane + r * C;
  const float *cr = cost + p * C;
  float z[LPL];
  float mx = -__builtin_inff();
#pragma unroll
  for (int k = 0; k < LPL; ++k) {
    const int c = lane + 64 * k;
    if (c < C) {
      const float e = (((m0[c] + m0[plane + c]) + m0[2 * plane + c]) + m0[3 * plane + c]) + cr[c];
      z[k] = -e * inv_T;
    } else {
      z[k] = -__builtin_inff();
    }
    mx = fmaxf(mx, z[k]);
  }
  mx = seg_reduce<64>(mx, OpMax());
  float s = 0.f;
  float b[LPL];
#pragma unroll
  for (int k = 0; k < LPL; ++k) {
    const int c = lane + 64 * k;
    b[k] = c < C ? fast_exp(z[k] - mx) : 0.f;
    s += b[k];
  }
  s = seg_reduce<64>(s, OpSum());
  const float rs = 1.f / s;
  float bv = -1.f;
  int bk = 0;
#pragma unroll
  for (int k = 0; k < LPL; ++k) {
    const int c = lane + 64 * k;
    b[k] *= rs;
    if (c < C && beliefs) beliefs[p * C + c] = b[k];
    // argmax with first-index ties: minimise (-belief, label) lexicographically
    if (c < C && (b[k] > bv)) { bv = b[k]; bk = c; }
  }
  if (label) {
    float nv = -bv;
    int key = -bk;  // larger key wins ties -> smaller label
    seg_argmin<64>(nv, key);
    if (lane == 0) label[p] = (float)(-key);
  }
}

}  // namespace
}  // namespace bpl

using namespace bpl;

extern "C" int bpl_softmax_labels(const float *x, float *y, int64_t rows, int64_t C, float scale, float sign,
                                  void *stream) {
  BPL_REQUIRE(x && y, BPL_E_ARG, "bpl_softmax_labels: null pointer");
  BPL_REQUIRE(rows >= 1 && C >= 1, BPL_E_ARG, "bpl_softmax_labels: sizes must be positive");
  BPL_REQUIRE(C <= 64 * kMaxLPL, BPL_E_LIMIT, "bpl_softmax_labels: C=%lld exceeds 256", (long long)C);
  const unsigned grid = (unsigned)((rows + 3) / 4);
  hipStream_t st = as_stream(stream);
  const int lpl = (int)((C + 63) / 64);
  KernelTimer kt(BPL_KERNEL_SOFTMAX, st);
  switch (lpl) {
    case 1: hipLaunchKernelGGL(softmax_labels_k<1>, dim3(grid), dim3(256), 0, st, x, y, rows, (int)C, scale, sign); break;
    case 2: hipLaunchKernelGGL(softmax_labels_k<2>, dim3(grid), dim3(256), 0, st, x, y, rows, (int)C, scale, sign); break;
    case 3: hipLaunchKernelGGL(softmax_labels_k<3>, dim3(grid), dim3(256), 0, st, x, y, rows, (int)C, scale, sign); break;
    default: hipLaunchKernelGGL(softmax_labels_k<4>, dim3(grid), dim3(256), 0, st, x, y, rows, (int)C, scale, sign); break;
  }
  return check_hip(hipGetLastError(), "bpl_softmax_labels launch");
}

extern "C" int bpl_lbp_beliefs(const float *cost, const float *messages, float *beliefs, float *label, int64_t N,
                               int64_t H, int64_t W, int64_t C, float T, void *stream) {
  BPL_REQUIRE(cost && messages && (beliefs || label), BPL_E_ARG, "bpl_lbp_beliefs: null pointer");
  BPL_REQUIRE(N >= 1 && H >= 1 && W >= 1 && C >= 1, BPL_E_ARG, "bpl_lbp_beliefs: sizes must be positive");
  BPL_REQUIRE(C <= 64 * kMaxLPL, BPL_E_LIMIT, "bpl_lbp_beliefs: C=%lld exceeds 256", (long long)C);
  BPL_REQUIRE(T > 0.f, BPL_E_ARG, "bpl_lbp_beliefs: temperature must be positive");
  const int64_t px = N * H * W;
  const unsigned grid = (unsigned)((px + 3) / 4);
  hipStream_t st = as_stream(stream);
  const int lpl = (int)((C + 63) / 64);
  const float it = 1.f / T;
  KernelTimer kt(BPL_KERNEL_BELIEFS, st);
  switch (lpl) {
    case 1: hipLaunchKernelGGL(beliefs_k<1>, dim3(grid), dim3(256), 0, st, cost, messages, beliefs, label, (int)N, (int)H, (int)W, (int)C, it); break;
    case 2: hipLaunchKernelGGL(beliefs_k<2>, dim3(grid), dim3(256), 0, st, cost, messages, beliefs, label, (int)N, (int)H, (int)W, (int)C, it); break;
    case 3: hipLaunchKernelGGL(beliefs_k<3>, dim3(grid), dim3(256), 0, st, cost, messages, beliefs, label, (int)N, (int)H, (int)W, (int)C, it); break;
    default: hipLaunchKernelGGL(beliefs_k<4>, dim3(grid), dim3(256), 0, st, cost, messages, beliefs, label, (int)N, (int)H, (int)W, (int)C, it); break;
  }
  return check_hip(hipGetLastError(), "bpl_lbp_beliefs launch");
}
