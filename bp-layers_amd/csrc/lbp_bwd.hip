// Min-sum LBP backward on gfx950: adjoint of one iteration, replayed from stored argmins.
//
// Semantics follow the reference backward (lbp_cuda_backward_kernel_reduction_min_sum,
// ops/lbp_stereo/src/lbp_min_sum_kernel.cu:215-417; host loop :513-621 and the semantic
// twins); the design is our own and deterministic (no float atomics):
//   phase V  adjoints of the DOWN and UP passes, one wave per (column, pass), walking the
//            column in reverse pass order; writes the h-gradient of every sender (X_D, X_U);
//   phase H  adjoints of the RIGHT and LEFT passes; the gradient reaching a written
//            message adds the vertical h-gradients of the receiver (X_D + X_U) as the
//            cross term (un-normalised, as in the reference: it sums to zero);
//   combine  grad_cost = X_D + X_U + X_R + X_L and, for the exact mode, the gradient of
//            the iteration's input messages (including the never-written boundary slots).
// Routing a receiver-label gradient g(c) to the sender label argmin(c) is a gather over
// the few label offsets that occur (|argmin(c) - c| <= D) plus one segment sum for the
// labels whose value came from the truncation term (all routed to the global argmin).
#include "bpl_common.h"

namespace bpl {
namespace {

struct BwdArgs {
  const float *jump, *edge, *in_grad, *msgs;
  const uint8_t *argmin, *gargmin;
  float *grad_jump, *grad_edge;
  float *X[4];        // per-direction sender h-gradients [N,H,W,C], indexed by pass direction
  float *partial;     // semantic global: per-(wave, segment) C x C partial sums
  int N, H, W, C, K, Nj, mode, n_base, accum, mat_global;
};

__device__ __forceinline__ int64_t px(int n, int y, int x, int H, int W) {
  return ((int64_t)n * H + y) * (int64_t)W + x;
}
__device__ __forceinline__ int64_t sl_px(int n, int s, int y, int x, int H, int W) {
  return (((int64_t)n * 4 + s) * H + y) * (int64_t)W + x;
}

template <int KIND, int SEG, int LPL, bool VERT, int PF>
__global__ __launch_bounds__(64) void lbp_bwd_phase(BwdArgs a) {
  constexpr int SPW = kWave / SEG;
  constexpr int CP = SEG * LPL;
  __shared__ float lds_g[2][SPW][CP];
  __shared__ int lds_a[2][SPW][CP];
  __shared__ float lds_p[2][SPW][16];

  const int H = a.H, W = a.W, C = a.C, K = a.K;
  const int L = VERT ? W : H;
  const int len = VERT ? H : W;
  const int nsteps = len - 1;
  const int wpg = (L + SPW - 1) / SPW;
  const int g_ = blockIdx.x / wpg;
  const int wl = blockIdx.x - g_ * wpg;
  const int n = a.n_base + (g_ >> 1);
  const int dsel = g_ & 1;
  const int dir = VERT ? (dsel ? UP : DOWN) : (dsel ? LEFT : RIGHT);
  const int rslot = dir == RIGHT ? LEFT : dir == LEFT ? RIGHT : dir == DOWN ? UP : DOWN;
  const bool fwd = (dir == RIGHT || dir == DOWN);
  const int lane = threadIdx.x;
  const int seg = lane / SEG;
  const int sl = lane - seg * SEG;
  const int line = wl * SPW + seg;
  const bool line_ok = line < L;
  const int ln = line_ok ? line : 0;
  const int nj = a.Nj > 1 ? n : 0;
  const bool exact = a.mode == BPL_GRAD_EXACT;
  float *X = a.X[dir];
  float *part = nullptr;
  if (KIND == BPL_LBP_SEMANTIC_GLOBAL) part = a.partial + ((int64_t)blockIdx.x * SPW + seg) * C * C;

  auto pos_of = [&](int s) { return fwd ? s : len - 1 - s; };
  auto yx_of = [&](int p, int &y, int &x) {
    if (VERT) { y = p; x = ln; } else { y = ln; x = p; }
  };
  const int64_t HW = (int64_t)H * W;

  // ---- the last position of the pass sends nothing: zero its outputs
  {
    int y, x;
    yx_of(pos_of(nsteps), y, x);
    const int64_t o = px(n, y, x, H, W);
#pragma unroll
    for (int k = 0; k < LPL; ++k) {
      const int c = sl + SEG * k;
      if (line_ok && c < C) X[o * C + c] = 0.f;
    }
    if (line_ok) {
      if (sl == 0) a.grad_edge[sl_px(n, dir, y, x, H, W)] = 0.f;
      if (KIND == BPL_LBP_STEREO && sl < K && !a.accum)
        a.grad_jump[(((int64_t)nj * 4 + dir) * K + sl) * HW + (int64_t)y * W + x] = 0.f;
      if (KIND == BPL_LBP_SEMANTIC_PIXEL && !a.accum) {
        float *G = a.grad_jump + px(dir, y, x, H, W) * C * C;
        for (int i = sl; i < C * C; i += SEG) G[i] = 0.f;
      }
    }
  }
  if (KIND == BPL_LBP_SEMANTIC_GLOBAL) {
    for (int i = sl; i < C * C; i += SEG) part[i] = 0.f;
    __syncthreads();  // the entries are later updated by their column's lane
  }

  // ---- prefetch ring: in_grad, messages, argmins and the cross term at the receiver
  float pg[PF][LPL], pm[PF][LPL], pd[PF][LPL], pu[PF][LPL], pp[PF];
  int pa[PF][LPL], pl[PF];
  auto load_step = [&](int s, float *dg, float *dm, int *da, float *dd, float *du, float &dp, int &dl) {
    int y, x;
    yx_of(pos_of(s), y, x);
    int yq = y, xq = x;
    if (VERT) yq += fwd ? 1 : -1; else xq += fwd ? 1 : -1;
    const int64_t oq = sl_px(n, rslot, yq, xq, H, W);
    const int64_t pq = px(n, yq, xq, H, W) * C;
#pragma unroll
    for (int k = 0; k < LPL; ++k) {
      const int c = min(sl + SEG * k, C - 1);  // clamped: loads stay unconditional
      dg[k] = a.in_grad[oq * C + c];
      dm[k] = a.msgs[oq * C + c];
      da[k] = (int)a.argmin[oq * C + c];
      if (!VERT) {
        dd[k] = a.X[DOWN][pq + c];
        du[k] = a.X[UP][pq + c];
      } else {
        dd[k] = 0.f;
        du[k] = 0.f;
      }
    }
    dl = (int)a.gargmin[oq];
    const int64_t eo = sl_px(n, dir, y, x, H, W);
    const int j = (KIND == BPL_LBP_STEREO && sl < K) ? sl : -1;
    dp = j >= 0 ? a.jump[(((int64_t)nj * 4 + dir) * K + j) * HW + (int64_t)y * W + x] : a.edge[eo];
  };

  if (nsteps > 0) {  // a single-pixel scanline has no receiver to read
#pragma unroll
    for (int i = 0; i < PF; ++i) {
      const int s = nsteps - 1 - i;
      load_step(s >= 0 ? s : 0, pg[i], pm[i], pa[i], pd[i], pu[i], pp[i], pl[i]);
      // keep the slots in issue order: the loop's wait counts assume slot 0 is the oldest
      __builtin_amdgcn_sched_barrier(0);
    }
  }

  float acc[LPL];
#pragma unroll
  for (int k = 0; k < LPL; ++k) acc[k] = 0.f;

  for (int t0 = 0; t0 < nsteps; t0 += PF) {
#pragma unroll
    for (int i = 0; i < PF; ++i) {
      const int t = t0 + i;
      const int s = nsteps - 1 - t;  // adjoint order
      if (t < nsteps) {
        float *G = pg[i], *Mq = pm[i], *XD = pd[i], *XU = pu[i];
        int *A = pa[i];
        const float pv = pp[i];
        const int gl = pl[i];
        const int buf = t & 1;
        int y, x;
        yx_of(pos_of(s), y, x);

        // normalisation Jacobian: g = t - softmax(msg) * sum(t), t = in_grad + chain
        float tt[LPL], ts = 0.f;
#pragma unroll
        for (int k = 0; k < LPL; ++k) {
          tt[k] = G[k] + acc[k];
          if (sl + SEG * k < C) ts += tt[k];
        }
        ts = seg_reduce<SEG>(ts, OpSum());
        float gv[LPL];
        float dmax = 0.f, fsum = 0.f;
#pragma unroll
        for (int k = 0; k < LPL; ++k) {
          const int c = sl + SEG * k;
          const bool ok = c < C;
          float gk = tt[k] - fast_exp(Mq[k]) * ts;
          if (!VERT) gk += XD[k] + XU[k];
          gk = ok ? gk : 0.f;
          gv[k] = gk;
          const bool mk = A[k] == kFromTrunc;
          if (ok) {
            lds_g[buf][seg][c] = gk;
            lds_a[buf][seg][c] = mk ? -1 : A[k];
            if (mk) fsum += gk;
            else dmax = fmaxf(dmax, (float)abs(A[k] - c));
          }
        }
        if (sl < 16) lds_p[buf][seg][sl] = pv;
        fsum = seg_reduce<SEG>(fsum, OpSum());
        dmax = seg_reduce<64>(dmax, OpMax());
        __syncthreads();
        const float w = lds_p[buf][seg][KIND == BPL_LBP_STEREO ? K : 0];

        // ---- route to the sender labels (deterministic gather)
        const int D = (int)dmax;
        float gh[LPL];
#pragma unroll
        for (int k = 0; k < LPL; ++k) gh[k] = 0.f;
        for (int kk = -D; kk <= D; ++kk) {
#pragma unroll
          for (int k = 0; k < LPL; ++k) {
            const int l = sl + SEG * k;
            const int c = l - kk;
            if (l < C && c >= 0 && c < C && lds_a[buf][seg][c] == l) gh[k] += lds_g[buf][seg][c];
          }
        }
#pragma unroll
        for (int k = 0; k < LPL; ++k)
          if (sl + SEG * k == gl) gh[k] += fsum;

        // ---- parameter gradients of (dir, sender)
        const int64_t eo = sl_px(n, dir, y, x, H, W);
        if (KIND == BPL_LBP_STEREO) {
          float ge = 0.f;
          for (int b = 1; b < K; ++b) {
            float v = 0.f;
#pragma unroll
            for (int k = 0; k < LPL; ++k) {
              const int c = sl + SEG * k;
              if (c < C) {
                const bool mk = A[k] == kFromTrunc;
                const int src = mk ? gl : A[k];
                const int bin = (mk && exact) ? K - 1 : min(max(K / 2 + c - src, 1), K - 1);
                v += bin == b ? gv[k] : 0.f;
              }
            }
            v = seg_reduce<SEG>(v, OpSum());
            ge += lds_p[buf][seg][b] * v;
            if (line_ok && sl == b) {
              float *dst = a.grad_jump + (((int64_t)nj * 4 + dir) * K + b) * HW + (int64_t)y * W + x;
              *dst = a.accum ? *dst + w * v : w * v;
            }
          }
          if (line_ok && sl == 0) {
            a.grad_edge[eo] = ge;
            if (!a.accum) a.grad_jump[(((int64_t)nj * 4 + dir) * K) * HW + (int64_t)y * W + x] = 0.f;
          }
        } else if (KIND == BPL_LBP_SEMANTIC_GLOBAL) {
          // entry [r][c] of this direction's matrix; every truncation-term winner uses the
          // truncation entry [0][C-1] in the exact mode and [gargmin][c] in the reference mode
          const float *Md = a.jump + (int64_t)dir * C * C;
          float ge = 0.f;
#pragma unroll
          for (int k = 0; k < LPL; ++k) {
            const int c = sl + SEG * k;
            if (c < C) {
              const bool mk = A[k] == kFromTrunc;
              const bool to_trunc = mk && exact;
              const int r = mk ? gl : A[k];
              ge += (to_trunc ? Md[C - 1] : Md[r * C + c]) * gv[k];
              if (line_ok && !to_trunc) part[r * C + c] += w * gv[k];
              if (line_ok && exact && c == C - 1) part[C - 1] += w * fsum;  // lane owning column C-1
            }
          }
          ge = seg_reduce<SEG>(ge, OpSum());
          if (line_ok && sl == 0) a.grad_edge[eo] = ge;
        } else {
          float *Gp = a.grad_jump + px(dir, y, x, H, W) * C * C;
          // edge gradient: zeros in the reference mode (not computed by the reference build),
          // sum_c f(c, argmin(c)) g(c) in the exact mode
          float ge = 0.f;
          if (exact) {
            const float *Jp = a.jump + px(dir, y, x, H, W) * C * C;
#pragma unroll
            for (int k = 0; k < LPL; ++k) {
              const int c = sl + SEG * k;
              if (c < C) ge += (A[k] == kFromTrunc ? Jp[C - 1] : Jp[A[k] * C + c]) * gv[k];
            }
            ge = seg_reduce<SEG>(ge, OpSum());
          }
          if (line_ok && sl == 0) a.grad_edge[eo] = ge;
          if (line_ok) {
            for (int r = 0; r < C; ++r) {
#pragma unroll
              for (int k = 0; k < LPL; ++k) {
                const int c = sl + SEG * k;
                if (c < C) {
                  const bool mk = A[k] == kFromTrunc;
                  float v = 0.f;
                  if (!mk && A[k] == r) v = w * gv[k];
                  if (mk && !exact && gl == r) v = w * gv[k];
                  if (exact && r == 0 && c == C - 1) v += w * fsum;
                  float *dst = Gp + r * C + c;
                  *dst = a.accum ? *dst + v : v;
                }
              }
            }
          }
        }

        // ---- chain and sender h-gradient
        const int64_t op = px(n, y, x, H, W) * C;
#pragma unroll
        for (int k = 0; k < LPL; ++k) {
          acc[k] = gh[k];
          const int c = sl + SEG * k;
          if (line_ok && c < C) X[op + c] = gh[k];
        }
      }
      // refill ring slot i after its last use (no register copies, no drain of the ring)
      load_step(s - PF >= 0 ? s - PF : 0, pg[i], pm[i], pa[i], pd[i], pu[i], pp[i], pl[i]);
    }
  }
}

// grad_cost and the input-message gradient from the four sender h-gradients.
__global__ void lbp_bwd_combine(const float *__restrict__ XL, const float *__restrict__ XR,
                                const float *__restrict__ XU, const float *__restrict__ XD,
                                const float *__restrict__ in_grad, float *__restrict__ grad_cost,
                                float *__restrict__ grad_msg, int N, int H, int W, int C, int exact) {
  const int64_t total = (int64_t)N * H * W * C;
  const int64_t HWC = (int64_t)H * W * C;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const float xl = XL[i], xr = XR[i], xu = XU[i], xd = XD[i];
    grad_cost[i] = ((xd + xu) + xr) + xl;
    const int64_t n = i / HWC;
    const int64_t r = i - n * HWC;            // (y, x, c) inside the image
    const int64_t base = n * 4 * HWC + r;     // slot 0 of this element
    if (!exact) {
      grad_msg[base] = in_grad[base];
      grad_msg[base + HWC] = in_grad[base + HWC];
      grad_msg[base + 2 * HWC] = in_grad[base + 2 * HWC];
      grad_msg[base + 3 * HWC] = in_grad[base + 3 * HWC];
      continue;
    }
    const int64_t yx = r / C;
    const int y = (int)(yx / W), x = (int)(yx - (int64_t)y * W);
    const float hr = xr + xl;   // read of U/D inputs by the horizontal passes
    const float sv = xd + xu;   // read of L/R boundary inputs by the vertical passes
    grad_msg[base + LEFT * HWC] = x == 0 ? (in_grad[base + LEFT * HWC] + xr) + sv : 0.f;
    grad_msg[base + RIGHT * HWC] = x == W - 1 ? (in_grad[base + RIGHT * HWC] + xl) + sv : 0.f;
    grad_msg[base + UP * HWC] = y == 0 ? (in_grad[base + UP * HWC] + xd) + hr : hr;
    grad_msg[base + DOWN * HWC] = y == H - 1 ? (in_grad[base + DOWN * HWC] + xu) + hr : hr;
  }
}

// Sum of the per-(wave, segment) C x C partials of each direction (fixed order).
__global__ void lbp_bwd_reduce_partials(const float *__restrict__ partial, float *__restrict__ out,
                                        int C, int nparts_per_dir_v, int nparts_per_dir_h,
                                        int64_t v_base, int64_t h_base, int N, int wpg_v, int wpg_h, int spw) {
  // partials are laid out in launch order: phase V then phase H; inside a phase block b
  // handles (group g = b / wpg, with direction bit g & 1), segments contiguous.
  const int d = blockIdx.y;
  const bool vert = d == UP || d == DOWN;
  const int dsel = vert ? (d == UP) : (d == LEFT);
  const int wpg = vert ? wpg_v : wpg_h;
  const int64_t base = vert ? v_base : h_base;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < C * C; i += gridDim.x * blockDim.x) {
    float s = 0.f;
    for (int nn = 0; nn < N; ++nn) {
      const int g = nn * 2 + dsel;
      for (int b = g * wpg; b < (g + 1) * wpg; ++b)
        for (int sg = 0; sg < spw; ++sg) s += partial[base + ((int64_t)b * spw + sg) * C * C + i];
    }
    out[(int64_t)d * C * C + i] = s;
  }
}

template <int KIND, int SEG, int LPL>
void launch_phase(const BwdArgs &a, bool vert, int ngroups, hipStream_t st) {
  constexpr int PF = 4;
  constexpr int SPW = kWave / SEG;
  const int L = vert ? a.W : a.H;
  const int wpg = (L + SPW - 1) / SPW;
  const dim3 grid((unsigned)(ngroups * 2 * wpg));
  if (vert)
    hipLaunchKernelGGL((lbp_bwd_phase<KIND, SEG, LPL, true, PF>), grid, dim3(kWave), 0, st, a);
  else
    hipLaunchKernelGGL((lbp_bwd_phase<KIND, SEG, LPL, false, PF>), grid, dim3(kWave), 0, st, a);
}

template <int KIND>
void launch_kind(const BwdArgs &a, bool vert, int ngroups, hipStream_t st) {
  const int C = a.C;
  if (C <= 16) launch_phase<KIND, 16, 1>(a, vert, ngroups, st);
  else if (C <= 32) launch_phase<KIND, 32, 1>(a, vert, ngroups, st);
  else if (C <= 64) launch_phase<KIND, 64, 1>(a, vert, ngroups, st);
  else if (C <= 128) launch_phase<KIND, 64, 2>(a, vert, ngroups, st);
  else if (C <= 192) launch_phase<KIND, 64, 3>(a, vert, ngroups, st);
  else launch_phase<KIND, 64, 4>(a, vert, ngroups, st);
}

int seg_width(int C) { return C <= 16 ? 16 : C <= 32 ? 32 : 64; }

void ws_layout(int kind, int64_t N, int64_t H, int64_t W, int64_t C, int64_t *planes, int64_t *part_v,
               int64_t *part_h, int64_t *total) {
  const int64_t plane = N * H * W * C;
  *planes = 4 * plane;
  int64_t pv = 0, ph = 0;
  if (kind == BPL_LBP_SEMANTIC_GLOBAL) {
    const int spw = kWave / seg_width((int)C);
    const int64_t wv = N * 2 * ((W + spw - 1) / spw), wh = N * 2 * ((H + spw - 1) / spw);
    pv = wv * spw * C * C;
    ph = wh * spw * C * C;
  }
  *part_v = pv;
  *part_h = ph;
  *total = (*planes + pv + ph) * (int64_t)sizeof(float);
}

}  // namespace
}  // namespace bpl

using namespace bpl;

extern "C" int64_t bpl_lbp_bwd_workspace_size(int kind, int64_t N, int64_t H, int64_t W, int64_t C) {
  int64_t planes, pv, ph, total;
  ws_layout(kind, N, H, W, C, &planes, &pv, &ph, &total);
  return total;
}

extern "C" int bpl_lbp_bwd(int kind, int mode, const float *jump, int64_t jump_k, int64_t jump_n,
                           const float *edge, const float *in_grad, const float *messages,
                           const uint8_t *argmin, const uint8_t *gargmin,
                           float *grad_cost, float *grad_jump, float *grad_edge, float *grad_msg,
                           void *workspace, int64_t workspace_bytes,
                           int64_t N, int64_t H, int64_t W, int64_t C, int delta, void *stream) {
  (void)delta;
  BPL_REQUIRE(kind >= 0 && kind <= 2, BPL_E_ARG, "bpl_lbp_bwd: unknown kind %d", kind);
  BPL_REQUIRE(mode == BPL_GRAD_REFERENCE || mode == BPL_GRAD_EXACT, BPL_E_ARG, "bpl_lbp_bwd: unknown mode %d", mode);
  BPL_REQUIRE(N >= 1 && H >= 1 && W >= 1 && C >= 1, BPL_E_ARG, "bpl_lbp_bwd: sizes must be positive");
  BPL_REQUIRE(C <= 255, BPL_E_LIMIT, "bpl_lbp_bwd: C=%lld labels exceeds the supported 255", (long long)C);
  BPL_REQUIRE(edge && in_grad && messages && argmin && gargmin && grad_cost && grad_jump && grad_edge && grad_msg,
              BPL_E_ARG, "bpl_lbp_bwd: null pointer");
  BPL_REQUIRE(jump || (kind == BPL_LBP_SEMANTIC_PIXEL && mode == BPL_GRAD_REFERENCE), BPL_E_ARG,
              "bpl_lbp_bwd: jump required");
  if (kind == BPL_LBP_STEREO) {
    BPL_REQUIRE(jump_k >= 2 && jump_k <= 15, BPL_E_LIMIT, "bpl_lbp_bwd: stereo jump K=%lld not in [2,15]",
                (long long)jump_k);
    BPL_REQUIRE(jump_n == 1 || jump_n == N, BPL_E_ARG, "bpl_lbp_bwd: jump batch must be 1 or N");
  }
  int64_t planes, pv, ph, total;
  ws_layout(kind, N, H, W, C, &planes, &pv, &ph, &total);
  BPL_REQUIRE(workspace && workspace_bytes >= total, BPL_E_ARG,
              "bpl_lbp_bwd: workspace of %lld bytes is smaller than the %lld required",
              (long long)workspace_bytes, (long long)total);
  float *ws = static_cast<float *>(workspace);
  const int64_t plane = N * H * W * C;
  BwdArgs a;
  a.jump = jump; a.edge = edge; a.in_grad = in_grad; a.msgs = messages;
  a.argmin = argmin; a.gargmin = gargmin;
  a.grad_jump = grad_jump; a.grad_edge = grad_edge;
  for (int d = 0; d < 4; ++d) a.X[d] = ws + d * plane;
  a.partial = ws + planes;
  a.N = (int)N; a.H = (int)H; a.W = (int)W; a.C = (int)C;
  a.K = kind == BPL_LBP_STEREO ? (int)jump_k : 0;
  a.Nj = kind == BPL_LBP_STEREO ? (int)jump_n : 1;
  a.mode = mode; a.mat_global = 1;
  hipStream_t st = as_stream(stream);
  // parameter gradients shared across the batch accumulate over n in launch order
  const bool per_n = (kind == BPL_LBP_STEREO && a.Nj == 1 && N > 1) || (kind == BPL_LBP_SEMANTIC_PIXEL && N > 1);
  for (int phase = 0; phase < 2; ++phase) {
    const bool vert = phase == 0;
    if (kind == BPL_LBP_SEMANTIC_GLOBAL) a.partial = ws + planes + (vert ? 0 : pv);
    const int nb = per_n ? (int)N : 1;
    for (int b = 0; b < nb; ++b) {
      a.n_base = per_n ? b : 0;
      a.accum = per_n && b > 0;
      const int ng = per_n ? 1 : (int)N;
      switch (kind) {
        case BPL_LBP_STEREO: launch_kind<BPL_LBP_STEREO>(a, vert, ng, st); break;
        case BPL_LBP_SEMANTIC_GLOBAL: launch_kind<BPL_LBP_SEMANTIC_GLOBAL>(a, vert, ng, st); break;
        default: launch_kind<BPL_LBP_SEMANTIC_PIXEL>(a, vert, ng, st); break;
      }
      int rc = check_hip(hipGetLastError(), "bpl_lbp_bwd phase launch");
      if (rc) return rc;
    }
  }
  {
    const int64_t total_el = plane;
    const int threads = 256;
    int64_t blocks = (total_el + threads - 1) / threads;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(lbp_bwd_combine, dim3((unsigned)blocks), dim3(threads), 0, st, a.X[LEFT], a.X[RIGHT],
                       a.X[UP], a.X[DOWN], in_grad, grad_cost, grad_msg, (int)N, (int)H, (int)W, (int)C,
                       mode == BPL_GRAD_EXACT ? 1 : 0);
    int rc = check_hip(hipGetLastError(), "bpl_lbp_bwd combine launch");
    if (rc) return rc;
  }
  if (kind == BPL_LBP_SEMANTIC_GLOBAL) {
    const int spw = kWave / seg_width((int)C);
    const int wpg_v = (int)((W + spw - 1) / spw), wpg_h = (int)((H + spw - 1) / spw);
    hipLaunchKernelGGL(lbp_bwd_reduce_partials, dim3((unsigned)((C * C + 255) / 256), 4), dim3(256), 0, st,
                       ws + planes, grad_jump, (int)C, 0, 0, (int64_t)0, pv, (int)N, wpg_v, wpg_h, spw);
    int rc = check_hip(hipGetLastError(), "bpl_lbp_bwd reduce launch");
    if (rc) return rc;
  }
  return BPL_OK;
}
