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
#include <mutex>
#include <type_traits>

#include "bpl_common.h"

namespace bpl {

struct BwdArgs {
  const float *jump, *edge, *in_grad, *msgs;
  const uint8_t *argmin, *gargmin;
  float *grad_jump, *grad_edge;
  float *X[4];        // per-direction sender h-gradients [N,H,W,C], indexed by pass direction
  float *grad_cost, *grad_msg;  // outputs written by the paired horizontal phase
  float *partial;     // semantic global: per-(wave, segment) C x C partial sums
  int N, H, W, C, K, Nj, mode, n_base, accum, mat_global;
  int delta;  // the forward's window half-width, or -1 when the caller does not know it
  // chained exact iterations: in_grad's LEFT/RIGHT slots are zero away from the row ends
  // (igz), grad_msg's LEFT/RIGHT slots need writing only at the row ends (osp)
  int igz, osp;
  // chained U/D slots (BPL_GRAD_IN_UD_SHARED / BPL_GRAD_OUT_UD_SHARED): in_grad's DOWN slot is
  // valid at rows 0 and H-1 only, its other rows being the UP slot's (uds_in); grad_msg's
  // DOWN slot is written at rows 0 and H-1 only (uds_out)
  int uds_in, uds_out;
  int accum_cost;  // grad_cost is added to (BPL_GRAD_ACCUMULATE) by its final writers
};

// Paired horizontal launches (stereo, exact mode, fast step).  Their instantiations dominate
// the build, so they live in one translation unit per labels-per-lane count
// (lbp_bwd_pair<L>.hip defines BPL_BWD_PAIR_LPL = L and includes this file) and compile in
// parallel with this one.
template <int LPL>
void launch_bwd_pair(const BwdArgs &a, bool full, int fk, int ngroups, hipStream_t st);
// the unpaired horizontal launches (reference mode, or no pairing), in the same units
template <int LPL>
void launch_bwd_h_unpaired(const BwdArgs &a, bool full, int fk, int ngroups, hipStream_t st);

namespace {

__device__ __forceinline__ int64_t px(int n, int y, int x, int H, int W) {
  return ((int64_t)n * H + y) * (int64_t)W + x;
}
__device__ __forceinline__ int64_t sl_px(int n, int s, int y, int x, int H, int W) {
  return (((int64_t)n * 4 + s) * H + y) * (int64_t)W + x;
}

// Slot of in_grad read at receiver row y of a pass whose messages land in slot rslot: with
// BPL_GRAD_IN_UD_SHARED the DOWN slot holds its own values at rows 0 and H-1 only, its other
// rows are those of the UP slot (the U and D slots of an input-message gradient are equal away
// from those rows).
__device__ __forceinline__ int ig_slot(const BwdArgs &a, int rslot, int y) {
  return (a.uds_in && rslot == DOWN && y != 0 && y != a.H - 1) ? UP : rslot;
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
  // columns of a vertical launch go to the XCDs in contiguous ranges (shared cache lines)
  const int bi = VERT ? xcd_contiguous_block((int)blockIdx.x, (int)gridDim.x) : (int)blockIdx.x;
  const int g_ = bi / wpg;
  const int wl = bi - g_ * wpg;
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
  if (KIND == BPL_LBP_SEMANTIC_GLOBAL) part = a.partial + ((int64_t)bi * SPW + seg) * C * C;

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
    const int64_t og = sl_px(n, ig_slot(a, rslot, yq), yq, xq, H, W);
#pragma unroll
    for (int k = 0; k < LPL; ++k) {
      const int c = min(sl + SEG * k, C - 1);  // clamped: loads stay unconditional
      dg[k] = (!VERT && a.igz) ? 0.f : a.in_grad[og * C + c];
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

// ---------------------------------------------------------------------------------------
// Stereo backward for 16 < C <= 256 with K <= 8: one scanline per wave, label
// c = lane * LPL + k (the forward's layout).  Same adjoint as lbp_bwd_phase; per step:
//   * routing: a sender label l collects g(c) of the receiver labels c = l - 3 .. l + 3 whose
//     argmin is l, reading the neighbour lanes' g and argmin through DPP shifts (an LDS
//     gather over the widest offset only when some |argmin(c) - c| > 3);
//   * pairwise gradients: per-lane sums for the K - 1 bins and the edge term
//     sum_b J[b] v[b] (bin slot 0) reduced together by one reduce-scatter;
//   * no per-label control flow: labels past C carry g = 0 and argmin -1.
// ---------------------------------------------------------------------------------------
template <int LPL>
__device__ __forceinline__ void ld_lanes_f(const float *row, uint32_t voff, float *d) {
  const float *p = reinterpret_cast<const float *>(reinterpret_cast<const char *>(row) + voff);
#pragma unroll
  for (int k = 0; k < LPL; ++k) d[k] = p[k];
}
__device__ __forceinline__ float shr1_f(float v, float fill) {  // lane i <- lane i-1
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(fill), __float_as_int(v), 0x138, 0xF, 0xF, false));
}
__device__ __forceinline__ float shl1_f(float v, float fill) {  // lane i <- lane i+1
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(fill), __float_as_int(v), 0x130, 0xF, 0xF, false));
}
__device__ __forceinline__ int shr1_i(int v, int fill) {
  return __builtin_amdgcn_update_dpp(fill, v, 0x138, 0xF, 0xF, false);
}
__device__ __forceinline__ int shl1_i(int v, int fill) {
  return __builtin_amdgcn_update_dpp(fill, v, 0x130, 0xF, 0xF, false);
}

// Fast step (FASTK = K in {6, 8}; delta <= 3, no batch accumulation; used for every scanline whose
// window offsets are all zero, so every routed argmin lies within 3 labels of its receiver).
// The step has no branches, so the compiler can overlap its tail (pairwise bins, stores) with
// the dependent chain of the next step.  The normalisation sum ts = sum_c (G + acc) is carried
// as a scalar recursion instead of a reduction of the chain: routing moves every receiver
// gradient to exactly one sender label, so sum_c acc_t = sum_c g_{t-1}
// = ts_{t-1} (1 - sum_c exp(M_{t-1})) + sum_c X_{t-1}, and
// ts_t = ts_{t-1} + sum_c (G_t + X_{t-1} - ts_{t-1} exp(M_{t-1})) needs only loaded data and
// the previous ts.  The scattered outputs (bins, edge) are raw buffer stores whose unused
// lanes carry an out-of-range offset and are dropped by the hardware.

// PAIR (horizontal phase, exact mode): the adjoints of the RIGHT and LEFT passes of one row
// run as the two waves of one workgroup.  They sweep the row in opposite directions, so every
// pixel is visited first by one wave and later by the other; after a barrier at the middle of
// the sweep (step Tb), the second visitor reads the first visitor's h-gradient and writes the
// pixel's grad_cost = ((X_D + X_U) + X_R) + X_L and the U and D slots X_R + X_L of its
// input-message gradient itself (the first visitor writes the L and R slots, 0), instead of a
// separate combine pass over four planes.  The
// pixels whose visits both fall before the barrier (the middle of the row) and the row ends
// are combined after a final barrier; the y-boundary terms of the U and D slots are added by
// lbp_bwd_pair_rows.
// HS (PAIR only, the chained-iteration flags as compile-time bits): bit 0 = in_grad's LEFT and
// RIGHT slots are zero inside the sweep (not loaded), bit 1 = grad_msg's LEFT and RIGHT slots
// are written at the row ends only.  The other kernels read the same flags from a.igz / a.osp.
// cache policy of the stereo backward's row stores (h-gradients, grad_cost, grad_msg):
// non-temporal (2), each plane being read back only after the phase has streamed its own inputs
// (same-run A/B at 384x1248x192: vertical 0.585 -> 0.569 ms, horizontal 0.932 -> 0.901 ms)
#ifndef BPL_BWD_ST_AUX
#define BPL_BWD_ST_AUX 2
#endif
// cache policy of its row loads: non-temporal (vertical 0.568 -> 0.553 ms, horizontal
// 0.900 -> 0.887 ms)
#ifndef BPL_BWD_LD_AUX
#define BPL_BWD_LD_AUX 2
#endif
// cache policy of its argmin loads: default (non-temporal: vertical 0.553 -> 0.563 ms)
#ifndef BPL_BWD_AM_AUX
#define BPL_BWD_AM_AUX 0
#endif
// cache policy of the semantic per-pixel kernels' matrix-row loads and matrix-gradient stores:
// default (non-temporal made every phase 8-10% slower: the two images of a wave and neighbouring
// scanlines share those lines)
#ifndef BPL_SEMP_AUX
#define BPL_SEMP_AUX 0
#endif
template <int LPL, bool VERT, bool FULL, int FASTK, bool PAIR, int PF, int HS = 0>
__global__ __launch_bounds__(PAIR ? 4 * kWave : kWave) void lbp_bwd_stereo_w64(BwdArgs a) {
  static_assert(!(PAIR && VERT), "pairing is for the horizontal phase");
  constexpr int ND = (3 + LPL - 1) / LPL;  // neighbour lanes reached by offsets up to 3
  constexpr int NWG = PAIR ? 4 : 1;        // waves per workgroup (PAIR: two rows, two passes each)
  __shared__ float lds_g_all[NWG][2][kWave * LPL];
  __shared__ int lds_a_all[NWG][2][kWave * LPL];

  const int H = a.H, W = a.W, C = a.C, K = a.K;
  const int L = VERT ? W : H;
  const int len = VERT ? H : W;
  const int nsteps = len - 1;
  // wave index in the workgroup, wave-uniform: everything derived from it (direction, start
  // pixel, stream pointers) then stays in scalar registers.  A paired workgroup holds two rows
  // (four waves: one wave per SIMD; two-wave workgroups put two waves on some SIMDs); with an
  // odd number of rows the last workgroup's second row repeats row L-1, whose stores then
  // write the same values twice (the paired kernel only stores, never accumulates).
  const int wq = PAIR ? __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)) : 0;
  const int wv = wq & 1;  // pass of the row (PAIR)
  const int L2 = (L + 1) / 2;
  // columns of a vertical launch go to the XCDs in contiguous ranges (shared cache lines)
  const int bi = VERT ? xcd_contiguous_block((int)blockIdx.x, (int)gridDim.x) : (int)blockIdx.x;
  const int g_ = PAIR ? 2 * (bi / L2) + wv : bi / L;
  const int ln = PAIR ? min(2 * (bi - (g_ >> 1) * L2) + (wq >> 1), L - 1) : bi - g_ * L;
  const int n = a.n_base + (g_ >> 1);
  const int dir = VERT ? ((g_ & 1) ? UP : DOWN) : ((g_ & 1) ? LEFT : RIGHT);
  const int rslot = dir == RIGHT ? LEFT : dir == LEFT ? RIGHT : dir == DOWN ? UP : DOWN;
  const bool fwd = (dir == RIGHT || dir == DOWN);
  const int lane = threadIdx.x & (kWave - 1);
  float(&lds_g)[2][kWave * LPL] = lds_g_all[wq];
  int(&lds_a)[2][kWave * LPL] = lds_a_all[wq];
  const int c0 = lane * LPL;
  const int nj = a.Nj > 1 ? n : 0;
  const bool exact = PAIR || a.mode == BPL_GRAD_EXACT;  // the paired kernel runs in exact mode only
  // in_grad's slot of a horizontal pass is zero at every receiver (which is never the pass's
  // first pixel), and grad_msg's LEFT/RIGHT slots are wanted at the row ends only
  const bool igz = !VERT && (PAIR ? (HS & 1) != 0 : a.igz != 0);
  constexpr bool kOsp = PAIR && (HS & 2) != 0;
  const int64_t HW = (int64_t)H * W;
  // sender pixel of forward step s: q0 + s * dq (in-plane index); its receiver is one further
  const int64_t dq = VERT ? (fwd ? (int64_t)W : -(int64_t)W) : (fwd ? 1 : -1);
  const int64_t q0 = VERT ? (int64_t)(fwd ? 0 : H - 1) * W + ln : (int64_t)ln * W + (fwd ? 0 : W - 1);
  const int64_t img = (int64_t)n * HW * C;
  const int64_t rsl = ((int64_t)n * 4 + rslot) * HW;
  float *Xo = a.X[dir] + img;
  const float *XDp = a.X[DOWN] + img;
  const float *XUp = a.X[UP] + img;
  const float *ig_p = a.in_grad + rsl * C;
  const float *m_p = a.msgs + rsl * C;
  const uint8_t *am_p = a.argmin + rsl * C;
  const uint8_t *gl_p = a.gargmin + rsl;
  const float *par_p = lane < K ? a.jump + (((int64_t)nj * 4 + dir) * K + lane) * HW
                                : a.edge + ((int64_t)n * 4 + dir) * HW;
  float *gj_p = a.grad_jump + ((int64_t)nj * 4 + dir) * K * HW;
  float *ge_p = a.grad_edge + ((int64_t)n * 4 + dir) * HW;

  uint32_t voff[LPL];
  bool valid[LPL];
#pragma unroll
  for (int k = 0; k < LPL; ++k) {
    valid[k] = FULL || c0 + k < C;
    voff[k] = (uint32_t)(FULL ? c0 + k : min(c0 + k, C - 1));
  }
  auto ld_row = [&](const float *row, float *d) {
    if (FULL) {
      ld_lanes_f<LPL>(row, voff[0] * 4u, d);
    } else {
#pragma unroll
      for (int k = 0; k < LPL; ++k) d[k] = row[voff[k]];
    }
  };
  // Kernels with the fast step (the host checks that every [H,W,C] plane is below 2^31 bytes)
  // address the float row streams through buffer resources: one wave-uniform byte offset per
  // row (q * C * 4, 32-bit) and the lane's constant byte offset, so no step forms per-lane
  // 64-bit addresses.
  constexpr bool BUF = FASTK > 0;
  const uint32_t plane_bytes = (uint32_t)(HW * C * 4);
  // BPL_GRAD_IN_UD_SHARED, UP pass (messages in the DOWN slot): rows 1 .. H-2 of in_grad come
  // from the UP slot, the plane before (ig_slot); the resource then covers both planes
  const bool uds = VERT && a.uds_in && rslot == DOWN;
  const int64_t HWCi = HW * C;
  auto ig_edge = [&](int64_t qr) { return qr < (int64_t)W || qr >= (int64_t)(H - 1) * W; };
  const __amdgpu_buffer_rsrc_t ig_rs = uds ? make_rsrc(ig_p - HWCi, 2u * plane_bytes) : make_rsrc(ig_p, plane_bytes);
  const __amdgpu_buffer_rsrc_t m_rs = make_rsrc(m_p, plane_bytes);
  const __amdgpu_buffer_rsrc_t xd_rs = make_rsrc(XDp, plane_bytes);
  const __amdgpu_buffer_rsrc_t xu_rs = make_rsrc(XUp, plane_bytes);
  const __amdgpu_buffer_rsrc_t am_rs = make_rsrc(am_p, plane_bytes / 4u);
  auto row_off = [&](int64_t q) -> uint32_t { return (uint32_t)q * (uint32_t)C * 4u; };
  auto ld_row_b = [&](__amdgpu_buffer_rsrc_t r, uint32_t soff, float *d) {
    if (FULL) {
      buf_ld_row<LPL, BPL_BWD_LD_AUX>(r, voff[0] * 4u, soff, d);
    } else {
#pragma unroll
      for (int k = 0; k < LPL; ++k) d[k] = buf_ld<BPL_BWD_LD_AUX>(r, voff[k] * 4u, soff);
    }
  };
  auto st_row_b = [&](__amdgpu_buffer_rsrc_t r, uint32_t soff, const float *v) {
    if (FULL) {
      buf_st_row<LPL, BPL_BWD_ST_AUX>(r, (uint32_t)c0 * 4u, soff, v);
    } else {
#pragma unroll
      for (int k = 0; k < LPL; ++k)
        buf_st<BPL_BWD_ST_AUX>(r, valid[k] ? (uint32_t)(c0 + k) * 4u : kBufDrop, soff, v[k]);
    }
  };

  // ---- the last position of the pass sends nothing: zero its outputs
  {
    const int64_t q = q0 + (int64_t)nsteps * dq;
#pragma unroll
    for (int k = 0; k < LPL; ++k)
      if (valid[k]) Xo[q * C + c0 + k] = 0.f;
    if (lane == 0) ge_p[q] = 0.f;
    if (lane < K && !a.accum) gj_p[(int64_t)lane * HW + q] = 0.f;
  }

  // argmin bytes stay in the ring as loaded (a full row of 2 or 3 bytes per lane arrives as
  // one 16-bit word plus one byte) and are split where they are used: a ring holding split
  // bytes makes the compiler split every slot at the loop latch, waiting for fresh loads
  constexpr int NRAW = FULL ? (LPL + 1) / 2 : LPL;
  auto ld_args = [&](const uint8_t *row, uint32_t *d) {
    if (FULL && LPL >= 2) {
      const uint8_t *p = row + voff[0];
      d[0] = (uint32_t)*reinterpret_cast<const uint16_t *>(p);
      if (LPL == 3) d[1] = (uint32_t)p[2];
    } else {
#pragma unroll
      for (int k = 0; k < NRAW; ++k) d[k] = (uint32_t)row[voff[k]];
    }
  };
  // the same through the buffer resource (byte offset soff of the row; the 16-bit access of
  // LPL = 3 is unaligned on odd lanes, as the global form above)
  auto ld_args_b = [&](uint32_t soff, uint32_t *d) {
    if (FULL && LPL >= 2) {
      d[0] = (uint32_t)__builtin_amdgcn_raw_buffer_load_b16(am_rs, voff[0], soff, BPL_BWD_AM_AUX);
      if (LPL == 3) d[1] = (uint32_t)__builtin_amdgcn_raw_buffer_load_b8(am_rs, voff[0] + 2u, soff, BPL_BWD_AM_AUX);
    } else {
#pragma unroll
      for (int k = 0; k < NRAW; ++k)
        d[k] = (uint32_t)__builtin_amdgcn_raw_buffer_load_b8(am_rs, voff[k], soff, BPL_BWD_AM_AUX);
    }
  };
  auto arg_of = [&](const uint32_t *d, int k) -> int {
    if (FULL && LPL >= 2) return k == 0 ? (int)(d[0] & 0xFFu) : k == 1 ? (int)(d[0] >> 8) : (int)d[1 < NRAW ? 1 : 0];
    return (int)d[k < NRAW ? k : 0];
  };

  float pg[PF][LPL], pm[PF][LPL], pd[PF][LPL], pu[PF][LPL], pp[PF];
  uint32_t pa[PF][NRAW];
  int pl[PF];
  auto load_step = [&](int s, float *dg, float *dm, uint32_t *da, float *dd, float *du, float &dp, int &dl) {
    // BUF: in-plane indices fit in 32 bits (the host's plane-size check)
    const int64_t q = BUF ? (int64_t)((int)q0 + s * (int)dq) : q0 + (int64_t)s * dq;
    const int64_t qr = BUF ? (int64_t)((int)q + (int)dq) : q + dq;
    const uint32_t so = row_off(qr);
    if (igz) {
#pragma unroll
      for (int k = 0; k < LPL; ++k) dg[k] = 0.f;
    } else if (BUF) {
      ld_row_b(ig_rs, uds && ig_edge(qr) ? so + plane_bytes : so, dg);
    } else {
      ld_row(ig_p + qr * C - (uds && !ig_edge(qr) ? HWCi : 0), dg);
    }
    if (BUF) ld_row_b(m_rs, so, dm);
    else ld_row(m_p + qr * C, dm);
    if (BUF) ld_args_b(so / 4u, da);
    else ld_args(am_p + qr * C, da);
    if (!VERT) {
      if (BUF) {
        ld_row_b(xd_rs, so, dd);
        ld_row_b(xu_rs, so, du);
      } else {
        ld_row(XDp + qr * C, dd);
        ld_row(XUp + qr * C, du);
      }
    } else {
#pragma unroll
      for (int k = 0; k < LPL; ++k) { dd[k] = 0.f; du[k] = 0.f; }
    }
    if constexpr (BUF) {
      // the receiver's global-argmin byte is wave-uniform: a scalar load of its aligned dword
      // (read-only data; counted by lgkmcnt, so it takes no slot of the vector-memory counter)
      const uintptr_t ga = reinterpret_cast<uintptr_t>(gl_p) + (uintptr_t)qr;
      const uint32_t word = *reinterpret_cast<const __attribute__((address_space(4))) uint32_t *>(ga & ~(uintptr_t)3);
      dl = (int)((word >> ((uint32_t)(ga & 3) * 8u)) & 0xFFu);
    } else {
      dl = (int)gl_p[qr];
    }
    dp = par_p[q];
  };
  if (nsteps > 0) {
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

  bool fast_done = false;           // PAIR: the fast sweep handled the row
  int fix_lo = 0, fix_hi = W - 1;   // PAIR: pixels combined after the sweep (plus the row ends)
  if constexpr (FASTK > 0) {
    constexpr int KK = FASTK;  // jump length of the fast step (K = a.K at launch)
    bool off = false;  // any non-zero window offset on this scanline (jump entry 0)
    const float *o_p = a.jump + ((int64_t)nj * 4 + dir) * K * HW;
    for (int p = lane; p < len; p += kWave) off = off || o_p[VERT ? (int64_t)p * W + ln : (int64_t)ln * W + p] != 0.f;
    bool fast = __ballot(off) == 0;
    if constexpr (PAIR) {  // all four waves (both passes of both rows) take the same path,
      __shared__ int pair_fast[4];  // so that they reach the same barriers
      if (lane == 0) pair_fast[wq] = fast ? 1 : 0;
      __syncthreads();
      fast = pair_fast[0] != 0 && pair_fast[1] != 0 && pair_fast[2] != 0 && pair_fast[3] != 0;
    }
    if (fast) {
      const int hw4 = __builtin_amdgcn_readfirstlane((int)(HW * 4));
      const __amdgpu_buffer_rsrc_t gj_rs = __builtin_amdgcn_make_buffer_rsrc(uniform_ptr(gj_p), (short)0, KK * hw4, 0x00020000);
      const __amdgpu_buffer_rsrc_t ge_rs = __builtin_amdgcn_make_buffer_rsrc(uniform_ptr(ge_p), (short)0, hw4, 0x00020000);
      const int row16 = lane >> 4;
      const int bin_here = 2 * row16 + (lane & 1);  // bin held by this lane after the reduce-scatter
      const bool bin_lane = (lane & 15) < 2 && bin_here < KK;
      constexpr int kDrop = 0x7FFFFFFF;              // out-of-range buffer offset: store dropped
      float gvprev[LPL];  // the previous step's h-gradient, before routing
#pragma unroll
      for (int k = 0; k < LPL; ++k) gvprev[k] = 0.f;
      // PAIR: the other pass's h-gradient at the sender, read by the second visitor
      const float *Xt_p = a.X[dir == RIGHT ? LEFT : RIGHT] + img;
      const bool isR = dir == RIGHT;
      float xo[PAIR ? PF : 1][LPL];
      float gco[PAIR ? PF : 1][LPL];  // PAIR, accumulating: grad_cost rows of the second visits
      const bool accg = PAIR && a.accum_cost != 0;
      // row streams of the fast step, buffer-addressed (see BUF)
      const int64_t HWC = HW * C;
      // a null grad_msg (not wanted): resources of 0 bytes, whose stores are dropped
      float *gm_p = a.grad_msg ? a.grad_msg + (int64_t)n * 4 * HWC : a.grad_cost;
      const uint32_t gm_bytes = a.grad_msg ? plane_bytes : 0u;
      const __amdgpu_buffer_rsrc_t xt_rs = make_rsrc(Xt_p, plane_bytes);
      const __amdgpu_buffer_rsrc_t xo_rs = make_rsrc(Xo, plane_bytes);
      const __amdgpu_buffer_rsrc_t gc_rs = make_rsrc(a.grad_cost + img, plane_bytes);
      const __amdgpu_buffer_rsrc_t gu_rs = make_rsrc(a.grad_msg ? gm_p + UP * HWC : gm_p, gm_bytes);
      // BPL_GRAD_OUT_UD_SHARED: the DOWN slot is written at rows 0 and H - 1 only
      const bool gd_row = !a.uds_out || ln == 0 || ln == H - 1;
      const __amdgpu_buffer_rsrc_t gd_rs = make_rsrc(a.grad_msg ? gm_p + DOWN * HWC : gm_p, gd_row ? gm_bytes : 0u);
      const __amdgpu_buffer_rsrc_t gl_rs = make_rsrc(a.grad_msg ? gm_p + LEFT * HWC : gm_p, gm_bytes);
      const __amdgpu_buffer_rsrc_t gr_rs = make_rsrc(a.grad_msg ? gm_p + RIGHT * HWC : gm_p, gm_bytes);
      // horizontal phase: the pairwise-bin and edge gradients of four consecutive steps (four
      // consecutive pixels of the row) are kept in registers and stored as one 16-byte access
      // per lane (fewer memory instructions per step leave room in the 6-bit vmcnt counter
      // for the prefetch ring); groups of PF steps start at multiples of 4
      static_assert(VERT || PF % 4 == 0, "horizontal groups hold whole blocks of four steps");
      float bq[4], eq[4];
      auto flush4 = [&](int tb) {  // steps tb .. tb + 3
        const int64_t sa = nsteps - 1 - tb;
        const int qa = (int)(q0 + sa * dq), qd = (int)(q0 + (sa - 3) * dq);
        const int qmin = min(qa, qd);
        float vb[4], ve[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {  // element k = pixel qmin + k (RIGHT: steps run right to left)
          vb[k] = fwd ? bq[3 - k] : bq[k];
          ve[k] = fwd ? eq[3 - k] : eq[k];
        }
        buf_st_row<4>(gj_rs, bin_lane ? (uint32_t)(bin_here * (int)HW + qmin) * 4u : (uint32_t)kDrop, 0u, vb);
        buf_st_row<4>(ge_rs, lane == 0 ? (uint32_t)qmin * 4u : (uint32_t)kDrop, 0u, ve);
      };
      auto step = [&](int i, int t, auto second_tag, auto batch_tag) {
        constexpr bool SECOND = decltype(second_tag)::value;
        constexpr bool BATCH = !VERT && decltype(batch_tag)::value;
        const int s = nsteps - 1 - t;
        const int64_t q = q0 + (int64_t)s * dq;
        // Uses of ring slot i wait for the previous step (no instruction; see the forward's
        // fast step): the work on a slot that does not depend on the chain (argmin bytes,
        // exp of the messages, X_D + X_U) would otherwise be hoisted to the top of the loop
        // body for every slot of the group, with waits for all of the group's loads there.
        // The argmin bytes, messages and vertical terms are ordered after the previous step's
        // h-gradient before routing, so that they overlap that step's routing.
#pragma unroll
        for (int r = 0; r < NRAW; ++r) asm("" : "+v"(pa[i][r]) : "v"(gvprev[0]));
#pragma unroll
        for (int k = 0; k < LPL; ++k) {
          asm("" : "+v"(pm[i][k]) : "v"(gvprev[0]));
          if constexpr (!VERT) asm("" : "+v"(pd[i][k]), "+v"(pu[i][k]) : "v"(gvprev[0]));
        }
        int A[LPL];
#pragma unroll
        for (int k = 0; k < LPL; ++k) A[k] = arg_of(pa[i], k);
        float pv = pp[i];
        asm("" : "+v"(pv) : "v"(acc[0]));
        int gl = pl[i];
        asm("" : "+v"(gl) : "v"(acc[0]));
        // normalisation sum sum_c (g + acc)(c): the routing moves every entry of the previous
        // step's h-gradient to one sender label, so sum acc = sum gvprev, and the reduction
        // runs beside the routing instead of after it.  Summing the actual vectors each step
        // projects out all the mass present, rounding included, as a direct sum does; a
        // recursion on the sums alone would let the rounding accumulate along the scanline.
        float r = 0.f;
#pragma unroll
        for (int k = 0; k < LPL; ++k) r += valid[k] ? pg[i][k] + gvprev[k] : 0.f;
        const float ts = wave_sum_f(r);
        float sm[LPL], X[LPL], gv[LPL], fs = 0.f;
        int ar[LPL];
#pragma unroll
        for (int k = 0; k < LPL; ++k) {
          sm[k] = valid[k] ? fast_exp(pm[i][k]) : 0.f;
          X[k] = (!VERT && valid[k]) ? pd[i][k] + pu[i][k] : 0.f;
          const float gk = ((pg[i][k] + acc[k]) - sm[k] * ts) + X[k];
          gv[k] = valid[k] ? gk : 0.f;
          const bool mk = A[k] == kFromTrunc;
          fs += mk ? gv[k] : 0.f;
          ar[k] = (valid[k] && !mk) ? A[k] : -1;
        }
        const float fsum = wave_sum_f(fs);
        // routing over offsets -3..3: gd[kk][k] = g of receiver c0 + k if its argmin is
        // c0 + k + kk, else 0; sender label c0 + j collects gd[kk] of receiver c0 + j - kk
        // (this lane or, through DPP shifts, the neighbour lanes), kk ascending
        float gd[7][LPL];
#pragma unroll
        for (int k = 0; k < LPL; ++k) {
          const int o = ar[k] >= 0 ? ar[k] - (c0 + k) : 99;
#pragma unroll
          for (int kk = -3; kk <= 3; ++kk) gd[kk + 3][k] = o == kk ? gv[k] : 0.f;
        }
        float gh[LPL];
#pragma unroll
        for (int k = 0; k < LPL; ++k) gh[k] = c0 + k == gl ? fsum : 0.f;
#pragma unroll
        for (int kk = -3; kk <= 3; ++kk) {
#pragma unroll
          for (int k = 0; k < LPL; ++k) {
            const int e = k - kk;
            float gc;
            if (e >= 0 && e < LPL) {
              gc = gd[kk + 3][e >= 0 && e < LPL ? e : 0];
            } else if (e < 0) {
              const int j = -e, d = (j + LPL - 1) / LPL, slot = d * LPL - j;
              gc = gd[kk + 3][slot];
#pragma unroll
              for (int m = 0; m < d; ++m) gc = shr1_f(gc, 0.f);
            } else {
              const int j = e - LPL + 1, d = (j + LPL - 1) / LPL, slot = j - 1 - (d - 1) * LPL;
              gc = gd[kk + 3][slot];
#pragma unroll
              for (int m = 0; m < d; ++m) gc = shl1_f(gc, 0.f);
            }
            gh[k] += gc;
          }
        }
#pragma unroll
        for (int k = 0; k < LPL; ++k) acc[k] = gh[k];

        // ---- tail: pairwise bins and edge term, stores
        const float w = readlane_f(pv, KK);
        float vb[8];
#pragma unroll
        for (int b = 0; b < 8; ++b) vb[b] = 0.f;
        // bin of a receiver routed by offset kk: clamp(K/2 - kk, 1, K-1); the truncation
        // winners go to bin K-1 (exact mode) or to the bin of their distance to gargmin
#pragma unroll
        for (int kk = -3; kk <= 3; ++kk) {
          const int b = min(max(KK / 2 - kk, 1), KK - 1);
          float sk = 0.f;
#pragma unroll
          for (int k = 0; k < LPL; ++k) sk += gd[kk + 3][k];
          vb[b] += sk;
        }
        if (exact) {
          vb[KK - 1] += fs;
        } else {
#pragma unroll
          for (int k = 0; k < LPL; ++k) {
            const bool mk = A[k] == kFromTrunc;
            const int bin = mk ? min(max(KK / 2 + c0 + k - gl, 1), KK - 1) : 0;
#pragma unroll
            for (int b = 1; b < KK; ++b) vb[b] += bin == b ? gv[k] : 0.f;
          }
        }
#pragma unroll
        for (int b = 1; b < KK; ++b) vb[0] = fmaf(readlane_f(pv, b), vb[b], vb[0]);
        float even, odd;
        reduce_scatter8(vb, even, odd);
        const float tot = (lane & 1) ? odd : even;
        const int qi = (int)q;
        const float bval = bin_here == 0 ? 0.f : w * tot;
        if constexpr (BATCH) {
          bq[i & 3] = bval;
          eq[i & 3] = tot;
        } else {
          __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(bval), gj_rs,
                                                bin_lane ? (bin_here * (int)HW + qi) * 4 : kDrop, 0, 0);
          __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(tot), ge_rs, lane == 0 ? qi * 4 : kDrop, 0, 0);
        }
        const uint32_t so = row_off(q);
        if constexpr (SECOND) {
          // second visitor: the pixel's grad_cost and the U and D slots of its input-message
          // gradient, X_R + X_L (X_D and X_U at the sender are the next step's receiver terms,
          // already in its ring slot)
          const int ni = (i + 1) % PF;
          float cst[LPL], hr[LPL];
#pragma unroll
          for (int k = 0; k < LPL; ++k) {
            const float xr = isR ? gh[k] : xo[i][k];
            const float xl = isR ? xo[i][k] : gh[k];
            float xdn = pd[ni][k], xun = pu[ni][k];   // ordered after this step's h-gradient
            asm("" : "+v"(xdn), "+v"(xun) : "v"(gv[0]));
            const float c = ((xdn + xun) + xr) + xl;
            cst[k] = accg ? gco[i][k] + c : c;
            hr[k] = xr + xl;
          }
          st_row_b(gc_rs, so, cst);
          st_row_b(gu_rs, so, hr);
          st_row_b(gd_rs, so, hr);
        } else {
          st_row_b(xo_rs, so, gh);
          if constexpr (PAIR && !kOsp) {
            // first visitor: the L and R slots, 0 away from the row ends (which are combined
            // after the sweep); written here so that both halves store three rows per step
            float zero[LPL];
#pragma unroll
            for (int k = 0; k < LPL; ++k) zero[k] = 0.f;
            st_row_b(gl_rs, so, zero);
            st_row_b(gr_rs, so, zero);
          }
        }
#pragma unroll
        for (int k = 0; k < LPL; ++k) gvprev[k] = gv[k];
      };
      auto load_xo = [&](int s, float *d, float *g) {
        const uint32_t so = row_off(q0 + (int64_t)s * dq);
        ld_row_b(xt_rs, so, d);
        if (accg) ld_row_b(gc_rs, so, g);
      };
      // a group of PF steps; each ring slot is refilled after its use (with the other pass's
      // h-gradient too in the second half)
      auto group = [&](int t0, auto second_tag) {
        constexpr bool SECOND = decltype(second_tag)::value;
#pragma unroll
        for (int i = 0; i < PF; ++i) {
          step(i, t0 + i, second_tag, std::true_type());
          if constexpr (!VERT) {
            if ((i & 3) == 3) flush4(t0 + i - 3);
          }
          const int s = nsteps - 1 - (t0 + i);
          const int sn = s - PF >= 0 ? s - PF : 0;
          load_step(sn, pg[i], pm[i], pa[i], pd[i], pu[i], pp[i], pl[i]);
          if constexpr (SECOND) load_xo(sn, xo[i], gco[i]);
        }
      };
      int t0 = 0;
      bool swept = false;
      if constexpr (PAIR) {
        // steps t < Tb (a multiple of PF, at least the middle of the row) write the pass's
        // h-gradient; steps Tb <= t < nsteps - 1 are second visits and combine.  The last
        // step (a row end) and the pixels of [W - 1 - Tb, Tb] are combined after the sweep.
        const int nl = nsteps - 1;
        const int Tb = PF * ((((W - 1) / 2) + PF - 1) / PF);
        if (Tb < nl) {
          for (; t0 < Tb; t0 += PF) group(t0, std::false_type());
          __syncthreads();  // the first half's h-gradients of both passes are written
#pragma unroll
          for (int i = 0; i < PF; ++i) {
            const int s = nsteps - 1 - (Tb + i);
            load_xo(s >= 0 ? s : 0, xo[i], gco[i]);
          }
          for (; t0 + PF <= nl; t0 += PF) group(t0, std::true_type());
#pragma unroll
          for (int i = 0; i < PF; ++i) {
            if (t0 + i < nl) step(i, t0 + i, std::true_type(), std::false_type());
            else if (t0 + i == nl) step(i, t0 + i, std::false_type(), std::false_type());
          }
          fix_lo = W - 1 - Tb;
          fix_hi = Tb;
          swept = true;
        }
      }
      if (!swept) {
        for (; t0 + PF <= nsteps; t0 += PF) group(t0, std::false_type());
#pragma unroll
        for (int i = 0; i < PF; ++i)
          if (t0 + i < nsteps) step(i, t0 + i, std::false_type(), std::false_type());
      }
      if constexpr (!PAIR) return;
      fast_done = true;
    }
  }

  if (!fast_done) {
    for (int t0 = 0; t0 < nsteps; t0 += PF) {
  #pragma unroll
      for (int i = 0; i < PF; ++i) {
        const int t = t0 + i;
        const int s = nsteps - 1 - t;  // adjoint order
        if (t < nsteps) {
          const int64_t q = q0 + (int64_t)s * dq;
          int A[LPL];
  #pragma unroll
          for (int k = 0; k < LPL; ++k) A[k] = arg_of(pa[i], k);
          const float pv = pp[i];
          const int gl = pl[i];

          // normalisation Jacobian: g = t - softmax(msg) * sum(t), t = in_grad + chain
          float tt[LPL], ts = 0.f;
  #pragma unroll
          for (int k = 0; k < LPL; ++k) {
            tt[k] = valid[k] ? pg[i][k] + acc[k] : 0.f;
            ts += tt[k];
          }
          ts = wave_sum_f(ts);
          float gv[LPL], fs = 0.f;
          bool wide = false;
  #pragma unroll
          for (int k = 0; k < LPL; ++k) {
            float gk = tt[k] - fast_exp(pm[i][k]) * ts;
            if (!VERT) gk += pd[i][k] + pu[i][k];
            gv[k] = valid[k] ? gk : 0.f;
            const bool mk = A[k] == kFromTrunc;
            fs += mk ? gv[k] : 0.f;
            wide = wide || (valid[k] && !mk && abs(A[k] - (c0 + k)) > 3);
          }
          const float fsum = wave_sum_f(fs);
          // argmin of the receiver labels, -1 where nothing is routed by offset
          int ar[LPL];
  #pragma unroll
          for (int k = 0; k < LPL; ++k) ar[k] = (valid[k] && A[k] != kFromTrunc) ? A[k] : -1;

          // ---- route to the sender labels (deterministic gather)
          float gh[LPL];
  #pragma unroll
          for (int k = 0; k < LPL; ++k) gh[k] = 0.f;
          if (__ballot(wide) == 0) {
            float gL[ND][LPL], gR[ND][LPL];
            int aL[ND][LPL], aR[ND][LPL];
  #pragma unroll
            for (int k = 0; k < LPL; ++k) {
              gL[0][k] = shr1_f(gv[k], 0.f);
              gR[0][k] = shl1_f(gv[k], 0.f);
              aL[0][k] = shr1_i(ar[k], -1);
              aR[0][k] = shl1_i(ar[k], -1);
            }
  #pragma unroll
            for (int d = 1; d < ND; ++d)
  #pragma unroll
              for (int k = 0; k < LPL; ++k) {
                gL[d][k] = shr1_f(gL[d - 1][k], 0.f);
                gR[d][k] = shl1_f(gR[d - 1][k], 0.f);
                aL[d][k] = shr1_i(aL[d - 1][k], -1);
                aR[d][k] = shl1_i(aR[d - 1][k], -1);
              }
  #pragma unroll
            for (int kk = -3; kk <= 3; ++kk) {
  #pragma unroll
              for (int k = 0; k < LPL; ++k) {
                const int e = k - kk;  // receiver c = l - kk, as a label offset from c0
                float gc;
                int ac;
                if (e >= 0 && e < LPL) {
                  gc = gv[e >= 0 && e < LPL ? e : 0];
                  ac = ar[e >= 0 && e < LPL ? e : 0];
                } else if (e < 0) {
                  const int j = -e, d = (j + LPL - 1) / LPL, slot = d * LPL - j;
                  gc = gL[d - 1][slot];
                  ac = aL[d - 1][slot];
                } else {
                  const int j = e - LPL + 1, d = (j + LPL - 1) / LPL, slot = j - 1 - (d - 1) * LPL;
                  gc = gR[d - 1][slot];
                  ac = aR[d - 1][slot];
                }
                gh[k] += ac == c0 + k ? gc : 0.f;
              }
            }
          } else {
            const int buf = t & 1;
            int dm = 0;
  #pragma unroll
            for (int k = 0; k < LPL; ++k) {
              lds_g[buf][c0 + k] = gv[k];
              lds_a[buf][c0 + k] = ar[k];
              if (ar[k] >= 0) dm = max(dm, abs(ar[k] - (c0 + k)));
            }
            const int D = wave_max_i(dm);
            wave_lds_sync();  // this wave's LDS rows only (a paired workgroup's other wave may not be here)
            for (int kk = -D; kk <= D; ++kk) {
  #pragma unroll
              for (int k = 0; k < LPL; ++k) {
                const int l = c0 + k;
                const int c = l - kk;
                if (c >= 0 && c < C && lds_a[buf][c] == l) gh[k] += lds_g[buf][c];
              }
            }
          }
  #pragma unroll
          for (int k = 0; k < LPL; ++k) gh[k] += c0 + k == gl ? fsum : 0.f;

          // ---- pairwise gradients of (dir, sender): bins 1..K-1 and the edge term in slot 0
          const float w = readlane_f(pv, K);
          float vb[8];
  #pragma unroll
          for (int b = 0; b < 8; ++b) vb[b] = 0.f;
  #pragma unroll
          for (int k = 0; k < LPL; ++k) {
            const bool mk = A[k] == kFromTrunc;
            const int src = mk ? gl : A[k];
            const int bin = (mk && exact) ? K - 1 : min(max(K / 2 + c0 + k - src, 1), K - 1);
  #pragma unroll
            for (int b = 1; b < 8; ++b) vb[b] += bin == b ? gv[k] : 0.f;
          }
  #pragma unroll
          for (int b = 1; b < 8; ++b)
            if (b < K) vb[0] = fmaf(readlane_f(pv, b), vb[b], vb[0]);
          float even, odd;
          reduce_scatter8(vb, even, odd);
          const int row = lane >> 4;
          const int bin_here = 2 * row + (lane & 1);
          const float tot = (lane & 1) ? odd : even;
          if ((lane & 15) < 2) {
            if (bin_here == 0) {
              ge_p[q] = tot;
              if (!a.accum) gj_p[q] = 0.f;
            } else if (bin_here < K) {
              float *dst = gj_p + (int64_t)bin_here * HW + q;
              *dst = a.accum ? *dst + w * tot : w * tot;
            }
          }

          // ---- chain and sender h-gradient
  #pragma unroll
          for (int k = 0; k < LPL; ++k) acc[k] = gh[k];
          if (FULL) {
            float *p = Xo + q * C + c0;
  #pragma unroll
            for (int k = 0; k < LPL; ++k) p[k] = gh[k];
          } else {
  #pragma unroll
            for (int k = 0; k < LPL; ++k)
              if (valid[k]) Xo[q * C + c0 + k] = gh[k];
          }
        }
        // refill ring slot i after its last use
        load_step(s - PF >= 0 ? s - PF : 0, pg[i], pm[i], pa[i], pd[i], pu[i], pp[i], pl[i]);
      }
    }
  }

  if constexpr (PAIR) {
    // combine the row ends and the pixels of [fix_lo, fix_hi] (every pixel when the sweep was
    // not paired); the U and D slots get X_R + X_L here, their y-boundary terms are added by
    // lbp_bwd_pair_rows
    __syncthreads();  // every h-gradient of both passes is written
    const int64_t HWC = HW * C;
    const float *XRp = a.X[RIGHT] + img, *XLp = a.X[LEFT] + img;
    const float *igr = a.in_grad + (int64_t)n * 4 * HWC;
    auto fix = [&](int x) {
      const int64_t q = (int64_t)ln * W + x;
      float xd[LPL], xu[LPL], xr[LPL], xl[LPL], ge[LPL], cst[LPL], hr[LPL], sl[LPL], sr[LPL];
      ld_row(XDp + q * C, xd);
      ld_row(XUp + q * C, xu);
      if (x <= W - 2) ld_row(XRp + q * C, xr);
      else
#pragma unroll
        for (int k = 0; k < LPL; ++k) xr[k] = 0.f;
      if (x >= 1) ld_row(XLp + q * C, xl);
      else
#pragma unroll
        for (int k = 0; k < LPL; ++k) xl[k] = 0.f;
#pragma unroll
      for (int k = 0; k < LPL; ++k) {
        const float sv = xd[k] + xu[k];
        cst[k] = (sv + xr[k]) + xl[k];
        hr[k] = xr[k] + xl[k];
        sl[k] = sr[k] = 0.f;
      }
      if (x == 0) {
        ld_row(igr + LEFT * HWC + q * C, ge);
#pragma unroll
        for (int k = 0; k < LPL; ++k) sl[k] = (ge[k] + xr[k]) + (xd[k] + xu[k]);
      }
      if (x == W - 1) {
        ld_row(igr + RIGHT * HWC + q * C, ge);
#pragma unroll
        for (int k = 0; k < LPL; ++k) sr[k] = (ge[k] + xl[k]) + (xd[k] + xu[k]);
      }
      float *gm = a.grad_msg ? a.grad_msg + (int64_t)n * 4 * HWC + q * C : nullptr;
      auto st = [&](float *row, const float *v) {
#pragma unroll
        for (int k = 0; k < LPL; ++k)
          if (valid[k]) row[c0 + k] = v[k];
      };
      if (a.accum_cost) {
        float old[LPL];
        ld_row(a.grad_cost + img + q * C, old);
#pragma unroll
        for (int k = 0; k < LPL; ++k) cst[k] = old[k] + cst[k];
      }
      st(a.grad_cost + img + q * C, cst);
      if (gm) {
        st(gm + UP * HWC, hr);
        if (!a.uds_out || ln == 0 || ln == H - 1) st(gm + DOWN * HWC, hr);
        if (!kOsp || x == 0) st(gm + LEFT * HWC, sl);
        if (!kOsp || x == W - 1) st(gm + RIGHT * HWC, sr);
      }
    };
    // the two waves take alternate pixels of the list (row ends first)
    int idx = 0;
    auto visit = [&](int x) {
      if ((idx++ & 1) == wv) fix(x);
    };
    visit(0);
    if (W > 1) visit(W - 1);
    for (int x = max(fix_lo, 1); x <= min(fix_hi, W - 2); ++x) visit(x);
  }
}

// ---------------------------------------------------------------------------------------
// Semantic global pairwise, 16 < C <= CM (20 or 32): the adjoint of one pass with
// lbp_bwd_phase's semantics, for either forward branch.  Two scanlines per wave, label
// c = lane & 31.
// Routing: receiver c writes its gradient g(c) into cell [r][c] of the segment's routing matrix
// in LDS (r its routing target: the argmin, or the label of the minimum for a truncation-term
// winner) after clearing the cell it wrote one step earlier; sender label l then sums row l in
// receiver-label order.  The row holds g(c) where c was routed to l and zero elsewhere, so the
// sum is the routed gradient with the additions of a label-order gather (bit for bit).
// The matrix gradient entry [r][c] += w g(c) (r the argmin; a truncation winner credits
// [0][C-1] in the exact mode, as a segment sum added by the lane owning column C-1, and [label
// of the minimum][c] in the reference mode) accumulates in an LDS copy of the segment's partial
// matrix; column c is touched only by lane c, so the sums are ordered and deterministic.  Its
// read and that of the matrix entry for the edge gradient are issued at the top of the step,
// off the step-to-step chain (in_grad + chain -> normalisation sum -> routing -> chain).  The
// partial is written once at the end for lbp_bwd_reduce_partials.
// TRUNC = false: the forward ran the full min-plus branch (delta >= C), so no argmin marks a
// truncation-term winner and the label-of-the-minimum bytes are not read.
// Planes are addressed through buffer resources (every [H,W,C] plane below 2^31 bytes) with
// scalar offsets that advance per step; lanes past C or past the last scanline store to an
// offset past the resource (dropped).
// ---------------------------------------------------------------------------------------
template <int CM, bool VERT, int PF, bool TRUNC>
__global__ __launch_bounds__(64) void lbp_bwd_semg(BwdArgs a) {
  constexpr int SEG = 32, SPW = 2;
  constexpr int RS = CM + 4;  // routing-matrix row stride (16-byte rows)
  __shared__ __align__(16) float lds_r[SPW][SEG * RS + SEG];  // rows 0..31, then a dummy row
  __shared__ float lds_mat[CM * CM];                  // the direction's matrix [r][c], stride CM
  __shared__ float lds_part[SPW][CM * CM + SEG];      // matrix-gradient partial [r][c], dummy slots
  const int H = a.H, W = a.W, C = a.C;
  const int L = VERT ? W : H;
  const int len = VERT ? H : W;
  const int nsteps = len - 1;
  const int wpg = (L + SPW - 1) / SPW;
  // columns of a vertical launch go to the XCDs in contiguous ranges (shared cache lines)
  const int bi = VERT ? xcd_contiguous_block((int)blockIdx.x, (int)gridDim.x) : (int)blockIdx.x;
  const int g_ = bi / wpg;
  const int wl = bi - g_ * wpg;
  const int n = a.n_base + (g_ >> 1);
  const int dir = VERT ? ((g_ & 1) ? UP : DOWN) : ((g_ & 1) ? LEFT : RIGHT);
  const int rslot = dir == RIGHT ? LEFT : dir == LEFT ? RIGHT : dir == DOWN ? UP : DOWN;
  const bool fwd = (dir == RIGHT || dir == DOWN);
  const int lane = threadIdx.x;
  const int seg = lane / SEG;
  const int c = lane - seg * SEG;
  const int line = wl * SPW + seg;
  const bool line_ok = line < L;
  const int ln = line_ok ? line : 0;
  const bool lab = c < C;
  const bool st_ok = line_ok && lab;
  const int cl = lab ? c : C - 1;
  const int64_t HW = (int64_t)H * W;
  const int64_t HWC = HW * C;
  const bool exact = a.mode == BPL_GRAD_EXACT;
  // pixel at position p of the scanline: lbase + p * pstride; wave-uniform byte offsets of a
  // position in planes of C floats, C bytes, one float and one byte per pixel
  const int lbase = VERT ? ln : ln * W;
  const uint32_t pstride = VERT ? (uint32_t)W : 1u;
  const uint32_t st_f = pstride * (uint32_t)C * 4u, st_b = pstride * (uint32_t)C, st_p = pstride * 4u;
  // one adjoint step moves the positions by -dpos (the pass runs forward in dpos)
  const uint32_t d_f = fwd ? st_f : 0u - st_f, d_b = fwd ? st_b : 0u - st_b;
  const uint32_t d_p = fwd ? st_p : 0u - st_p, d_1 = fwd ? pstride : 0u - pstride;
  auto pos = [&](int s) -> uint32_t { return (uint32_t)(fwd ? s : nsteps - s); };
  const uint32_t v_ld = (uint32_t)(lbase * C + cl) * 4u;
  const uint32_t v_ldb = (uint32_t)(lbase * C + cl);
  const uint32_t v_st = st_ok ? (uint32_t)(lbase * C + c) * 4u : kBufDrop;
  const uint32_t v_e = (uint32_t)lbase * 4u;
  const uint32_t v_pp = (line_ok && c == 0) ? (uint32_t)lbase * 4u : kBufDrop;
  const uint32_t v_p1 = (uint32_t)lbase;
  const uint32_t plane = (uint32_t)(HWC * 4);
  const int64_t img = (int64_t)n * HWC;
  const int64_t rsl = ((int64_t)n * 4 + rslot) * HW;
  // in_grad: with chained U/D slots the DOWN slot's rows 1..H-2 are the UP slot's, so the
  // resource spans [UP plane, DOWN plane] and only the row-0 receiver (the first adjoint step of
  // the UP pass) reads the DOWN plane
  const bool uds_dn = VERT && a.uds_in && rslot == DOWN;
  const auto ig_r = (!VERT && a.igz) ? make_rsrc(a.in_grad, 0u)
                  : uds_dn ? make_rsrc(a.in_grad + ((int64_t)n * 4 + UP) * HWC, 2u * plane)
                           : make_rsrc(a.in_grad + rsl * C, plane);
  const auto m_r = make_rsrc(a.msgs + rsl * C, plane);
  const auto am_r = make_rsrc(a.argmin + rsl * C, plane / 4u);
  const auto gl_r = make_rsrc(TRUNC ? a.gargmin + rsl : a.argmin, TRUNC ? (uint32_t)HW : 0u);
  const auto xd_r = make_rsrc(a.X[DOWN] + img, VERT ? 0u : plane);
  const auto xu_r = make_rsrc(a.X[UP] + img, VERT ? 0u : plane);
  const auto xo_r = make_rsrc(a.X[dir] + img, plane);
  const auto edge_r = make_rsrc(a.edge + ((int64_t)n * 4 + dir) * HW, (uint32_t)HW * 4u);
  const auto ge_r = make_rsrc(a.grad_edge + ((int64_t)n * 4 + dir) * HW, (uint32_t)HW * 4u);

  // the last position of the pass sends nothing: zero its outputs
  buf_st(xo_r, v_st, pos(nsteps) * st_f, 0.f);
  buf_st(ge_r, v_pp, pos(nsteps) * st_p, 0.f);
  const float *Md = a.jump + (int64_t)dir * C * C;
  for (int i = lane; i < CM * CM; i += kWave) {
    const int r = i / CM, cc = i - r * CM;
    lds_mat[i] = (r < C && cc < C) ? Md[r * C + cc] : 0.f;
  }
  for (int i = c; i < CM * CM + SEG; i += SEG) lds_part[seg][i] = 0.f;
  for (int i = c; i < SEG * RS + SEG; i += SEG) lds_r[seg][i] = 0.f;
  wave_lds_sync();

  float pg[PF], pm[PF], pd[PF], pu[PF], pw[PF];
  int pa[PF], pl[PF];
  // ring slot i of adjoint step s: the receiver's (position s + 1) in_grad, message, argmin,
  // label of the minimum and vertical h-gradients, the sender's (position s) edge weight
  auto load_at = [&](uint32_t rf, uint32_t rb, uint32_t r1, uint32_t qp, uint32_t igo, int i) {
    pg[i] = buf_ld(ig_r, v_ld, igo);
    pm[i] = buf_ld(m_r, v_ld, rf);
    pa[i] = (int)__builtin_amdgcn_raw_buffer_load_b8(am_r, v_ldb, rb, 0);
    pl[i] = TRUNC ? (int)__builtin_amdgcn_raw_buffer_load_b8(gl_r, v_p1, r1, 0) : 0;
    pd[i] = VERT ? 0.f : buf_ld(xd_r, v_ld, rf);
    pu[i] = VERT ? 0.f : buf_ld(xu_r, v_ld, rf);
    pw[i] = buf_ld(edge_r, v_e, qp);
  };
  auto load_step = [&](int s, int i) {
    const uint32_t pr = pos(s + 1);
    load_at(pr * st_f, pr * st_b, pr * pstride, pos(s) * st_p, pr * st_f + (uds_dn && pr == 0u ? plane : 0u), i);
  };
  if (nsteps > 0) {
#pragma unroll
    for (int i = 0; i < PF; ++i) {
      const int s = nsteps - 1 - i;
      load_step(s >= 0 ? s : 0, i);
      __builtin_amdgcn_sched_barrier(0);  // keep the ring slots in issue order
    }
  }
  float acc = 0.f;
  int prev = SEG * RS + c;  // the routing cell written one step earlier (the dummy row at first)
  // sender offsets of the current adjoint step
  uint32_t o_f = pos(nsteps - 1) * st_f, o_p = pos(nsteps - 1) * st_p;
  auto step = [&](int i) {
    const int A = pa[i];
    const bool tr = TRUNC && A == kFromTrunc;  // the value came from the truncation term
    const int r = tr ? pl[i] : A;              // routing target and reference-mode matrix row
    const bool to_tr = tr && exact;
    // LDS reads off the chain: the matrix entry of the edge gradient, the partial's entry
    // (24-bit multiplies: labels are below 256; the empty asm statements keep the values
    // computed on every lane, so the selects stay selects instead of branches)
    const int rcm = __mul24(r, CM);
    const float mv = to_tr ? lds_mat[C - 1] : lds_mat[rcm + cl];
    const int pidx = (st_ok && !to_tr) ? rcm + c : CM * CM + c;
    const float pold = lds_part[seg][pidx];
    // normalisation Jacobian: g = t - softmax(msg) * sum(t), t = in_grad + chain
    const float tt = pg[i] + acc;
    const float ts = seg_reduce<SEG>(lab ? tt : 0.f, OpSum());
    float gk = tt - fast_exp(pm[i]) * ts;
    if (!VERT) gk += pd[i] + pu[i];
    asm volatile("" : "+v"(gk));
    gk = lab ? gk : 0.f;
    // route: clear last step's cell, write [r][c]; sender c sums row c in label order
    int cell = __mul24(r, RS) + c;
    asm volatile("" : "+v"(cell));
    cell = lab ? cell : SEG * RS + c;
    lds_r[seg][prev] = 0.f;
    lds_r[seg][cell] = gk;
    prev = cell;
    wave_lds_sync();
    float row[CM];
#pragma unroll
    for (int l = 0; l < CM; ++l) row[l] = lds_r[seg][c * RS + l];
    wave_lds_sync();  // the row is read before the next step's writes
    float gh = 0.f;
#pragma unroll
    for (int l = 0; l < CM; ++l) gh += row[l];
    acc = gh;
    buf_st(xo_r, v_st, o_f, gh);
    // parameter gradients of (dir, sender): edge sum_c M[r][c] g(c) and matrix entry [r][c]
    const float w = pw[i];
    const float ge = seg_reduce<SEG>(lab ? mv * gk : 0.f, OpSum());
    buf_st(ge_r, v_pp, o_p, ge);
    lds_part[seg][pidx] = pold + w * gk;
    if constexpr (TRUNC) {
      if (exact) {
        const float fsum = seg_reduce<SEG>(lab && tr ? gk : 0.f, OpSum());
        if (line_ok && c == C - 1) lds_part[seg][C - 1] += w * fsum;
      }
    }
    o_f -= d_f;
    o_p -= d_p;
  };
  // main loop: the ring refills stay within the scanline (s - PF >= 0)
  int t0 = 0;
  {
    const int s = nsteps - 1 - PF;  // the first refill's sender position
    const uint32_t pr = pos(s >= 0 ? s + 1 : 0), pq = pos(s >= 0 ? s : 0);
    uint32_t l_f = pr * st_f, l_b = pr * st_b, l_1 = pr * pstride, l_p = pq * st_p;
    for (; t0 + 2 * PF <= nsteps; t0 += PF) {
#pragma unroll
      for (int i = 0; i < PF; ++i) {
        step(i);
        // refill ring slot i (its old contents are dead now: no register copies, no drain)
        load_at(l_f, l_b, l_1, l_p, l_f, i);
        l_f -= d_f;
        l_b -= d_b;
        l_1 -= d_1;
        l_p -= d_p;
      }
    }
  }
  // the last groups: refills clamped to the first sender, steps past the end skipped
  for (; t0 < nsteps; t0 += PF) {
#pragma unroll
    for (int i = 0; i < PF; ++i) {
      const int t = t0 + i;
      if (t < nsteps) {
        step(i);
        const int s = nsteps - 1 - t;
        load_step(s - PF >= 0 ? s - PF : 0, i);
      }
    }
  }
  // the segment's partial matrix, in lbp_bwd_reduce_partials' layout
  wave_lds_sync();
  float *part = a.partial + ((int64_t)bi * SPW + seg) * C * C;
  for (int i = c; i < C * C; i += SEG) {
    const int r = i / C, cc = i - r * C;
    part[i] = lds_part[seg][r * CM + cc];
  }
}

template <int CM>
void launch_semg(const BwdArgs &a, bool vert, int ngroups, hipStream_t st) {
  const int L = vert ? a.W : a.H;
  const dim3 grid((unsigned)(ngroups * 2 * ((L + 1) / 2)));
  // the forward's full min-plus branch (delta >= C) marks no truncation-term winners
  const bool trunc = !(a.delta >= a.C);
  if (vert && trunc) hipLaunchKernelGGL((lbp_bwd_semg<CM, true, 4, true>), grid, dim3(kWave), 0, st, a);
  else if (vert) hipLaunchKernelGGL((lbp_bwd_semg<CM, true, 4, false>), grid, dim3(kWave), 0, st, a);
  else if (trunc) hipLaunchKernelGGL((lbp_bwd_semg<CM, false, 4, true>), grid, dim3(kWave), 0, st, a);
  else hipLaunchKernelGGL((lbp_bwd_semg<CM, false, 4, false>), grid, dim3(kWave), 0, st, a);
}

// Semantic per-pixel pairwise, 16 < C <= 32 or C <= 20 (both forward branches, both modes):
// the adjoint step of lbp_bwd_semg (routing through the segment's sender x receiver matrix in
// LDS), with the per-pixel parameter gradients written per step: the sender's C x C matrix
// gradient (lane c writes column c: w g(c) at row argmin(c) -- row gargmin for a truncation
// winner in the reference mode; in the exact mode a truncation winner credits the truncation
// entry [0][C-1], which the lane owning column C-1 adds) and the edge gradient (sum_c
// M[r][c] g(c) in the exact mode from the matrix column read per step; zero in the reference
// mode, as the reference build).  IMGS (N == 2): the half-waves are the two images of the same
// scanline and the batch-shared matrix gradient is the sum image 0 + image 1 formed in
// registers (permlane32 swap), each half-wave writing half of the rows; otherwise (N == 1) two
// scanlines of one image, each writing its full slices.  Tensors are addressed as whole buffers
// (the [N,4,H,W,C] planes of all images below 2^31 bytes, a direction's matrix plane too) with
// per-lane offsets for the image, slot and matrix-row parts computed once and one scalar offset
// per position, so the steps carry no 64-bit address arithmetic; lanes that must not store get
// an offset past the buffer (dropped).
template <int CM, bool VERT, int PF, bool IMGS, bool EXACT>
__global__ __launch_bounds__(64) void lbp_bwd_semp(BwdArgs a) {
  constexpr int SEG = 32, SPW = 2;
  constexpr int RS = CM + 4;                          // routing-matrix row stride (16-byte rows)
  constexpr int NR = IMGS ? (CM + 1) / 2 : CM;        // matrix-gradient rows a lane writes per step
  __shared__ __align__(16) float lds_r[SPW][SEG * RS + SEG];  // rows 0..31, then a dummy row
  const int H = a.H, W = a.W, C = a.C;
  const int L = VERT ? W : H;
  const int len = VERT ? H : W;
  const int nsteps = len - 1;
  const int lane = threadIdx.x;
  const int seg = lane / SEG;
  const int c = lane - seg * SEG;
  const int wpg = IMGS ? L : (L + SPW - 1) / SPW;
  // columns of a vertical launch go to the XCDs in contiguous ranges (shared cache lines)
  const int bi = VERT ? xcd_contiguous_block((int)blockIdx.x, (int)gridDim.x) : (int)blockIdx.x;
  const int g_ = bi / wpg;
  const int wl = bi - g_ * wpg;
  const int n = IMGS ? seg : (g_ >> 1);
  const int dir = VERT ? ((g_ & 1) ? UP : DOWN) : ((g_ & 1) ? LEFT : RIGHT);
  const int rslot = dir == RIGHT ? LEFT : dir == LEFT ? RIGHT : dir == DOWN ? UP : DOWN;
  const bool fwd = (dir == RIGHT || dir == DOWN);
  const int line = IMGS ? wl : wl * SPW + seg;
  const bool line_ok = line < L;
  const int ln = line_ok ? line : 0;
  const bool lab = c < C;
  const bool st_ok = line_ok && lab;
  const int cl = lab ? c : C - 1;
  const uint32_t HWu = (uint32_t)H * (uint32_t)W, HWCu = HWu * (uint32_t)C;
  const uint32_t lbase = VERT ? (uint32_t)ln : (uint32_t)ln * (uint32_t)W;
  const uint32_t pstride = VERT ? (uint32_t)W : 1u;
  const uint32_t st_f = pstride * (uint32_t)C * 4u, st_b = pstride * (uint32_t)C, st_p = pstride * 4u;
  const uint32_t st_m = st_f * (uint32_t)C;
  // one adjoint step moves the positions by -dpos (the pass runs forward in dpos)
  const uint32_t d_f = fwd ? st_f : 0u - st_f, d_b = fwd ? st_b : 0u - st_b, d_p = fwd ? st_p : 0u - st_p;
  const uint32_t d_1 = fwd ? pstride : 0u - pstride, d_m = fwd ? st_m : 0u - st_m;
  auto pos = [&](int s) -> uint32_t { return (uint32_t)(fwd ? s : nsteps - s); };
  const uint32_t lpix = lbase * (uint32_t)C;
  const uint32_t N4 = (uint32_t)a.N * 4u;
  auto slot_f = [&](int sl, int lbl) { return (((uint32_t)n * 4u + (uint32_t)sl) * HWCu + lpix + (uint32_t)lbl) * 4u; };
  // in_grad: with chained U/D slots the DOWN slot's rows 1..H-2 are the UP slot's; only the
  // row-0 receiver (the first adjoint step of the UP pass) reads the DOWN slot (UP + 1)
  const bool uds_dn = VERT && a.uds_in && rslot == DOWN;
  const uint32_t v_ig = slot_f(uds_dn ? UP : rslot, cl);
  const uint32_t v_ms = slot_f(rslot, cl);
  const uint32_t v_am = ((uint32_t)n * 4u + (uint32_t)rslot) * HWCu + lpix + (uint32_t)cl;
  const uint32_t v_gl = ((uint32_t)n * 4u + (uint32_t)rslot) * HWu + lbase;
  const uint32_t v_x = ((uint32_t)n * HWCu + lpix + (uint32_t)cl) * 4u;
  const uint32_t v_xs = st_ok ? ((uint32_t)n * HWCu + lpix + (uint32_t)c) * 4u : kBufDrop;
  const uint32_t v_e = (((uint32_t)n * 4u + (uint32_t)dir) * HWu + lbase) * 4u;
  const uint32_t v_ge = (line_ok && c == 0) ? v_e : kBufDrop;
  const auto ig_r = (!VERT && a.igz) ? make_rsrc(a.in_grad, 0u) : make_rsrc(a.in_grad, N4 * HWCu * 4u);
  const auto m_r = make_rsrc(a.msgs, N4 * HWCu * 4u);
  const auto am_r = make_rsrc(a.argmin, N4 * HWCu);
  const auto gl_r = make_rsrc(a.gargmin, N4 * HWu);
  const auto xd_r = make_rsrc(a.X[DOWN], VERT ? 0u : (uint32_t)a.N * HWCu * 4u);
  const auto xu_r = make_rsrc(a.X[UP], VERT ? 0u : (uint32_t)a.N * HWCu * 4u);
  const auto xo_r = make_rsrc(a.X[dir], (uint32_t)a.N * HWCu * 4u);
  const auto edge_r = make_rsrc(a.edge, N4 * HWu * 4u);
  const auto ge_r = make_rsrc(a.grad_edge, N4 * HWu * 4u);
  const uint32_t mplane = HWCu * (uint32_t)C * 4u;
  const auto mat_r = make_rsrc(EXACT ? a.jump + (int64_t)dir * HWu * C * C : a.grad_jump, EXACT ? mplane : 0u);
  const auto gm_r = make_rsrc(a.grad_jump + (int64_t)dir * HWu * C * C, mplane);
  // matrix rows: row l of the sender's block, label cl (rows past C are never used); the
  // gradient rows this lane writes (IMGS: half of them per half-wave)
  const int rlo = IMGS ? (seg ? C / 2 : 0) : 0, rhi = IMGS ? (seg ? C : C / 2) : C;
  uint32_t v_mm[EXACT ? CM : 1], v_gm[NR];
#pragma unroll
  for (int l = 0; l < (EXACT ? CM : 1); ++l) v_mm[l] = ((lpix + (uint32_t)l) * (uint32_t)C + (uint32_t)cl) * 4u;
#pragma unroll
  for (int jr = 0; jr < NR; ++jr) {
    const int rr = rlo + jr;
    v_gm[jr] = (line_ok && lab && rr < rhi) ? ((lpix + (uint32_t)rr) * (uint32_t)C + (uint32_t)c) * 4u : kBufDrop;
  }

  // the last position of the pass sends nothing: zero its outputs (and its matrix gradient,
  // once per scanline)
  {
    const uint32_t pl = pos(nsteps);
    buf_st(xo_r, v_xs, pl * st_f, 0.f);
    buf_st(ge_r, v_ge, pl * st_p, 0.f);
    const bool zr = line_ok && lab && (!IMGS || seg == 0);
    for (int rr = 0; rr < C; ++rr)
      buf_st<BPL_SEMP_AUX>(gm_r, zr ? ((lpix + (uint32_t)rr) * (uint32_t)C + (uint32_t)c) * 4u : kBufDrop, pl * st_m, 0.f);
  }
  for (int i = c; i < SEG * RS + SEG; i += SEG) lds_r[seg][i] = 0.f;
  wave_lds_sync();

  constexpr int PJ = EXACT ? PF : 1;
  float pg[PF], pm[PF], pd[PF], pu[PF], pw[PF], pj[PJ][CM];
  int pa[PF], pl[PF];
  // ring slot i of adjoint step s: the receiver's (position s + 1) in_grad, message, argmin,
  // label of the minimum and vertical h-gradients, the sender's (position s) edge weight and
  // matrix column
  auto load_at = [&](uint32_t rf, uint32_t rb, uint32_t r1, uint32_t qp, uint32_t qm, uint32_t igo, int i) {
    pg[i] = buf_ld(ig_r, v_ig, igo);
    pm[i] = buf_ld(m_r, v_ms, rf);
    pa[i] = (int)__builtin_amdgcn_raw_buffer_load_b8(am_r, v_am, rb, 0);
    pl[i] = (int)__builtin_amdgcn_raw_buffer_load_b8(gl_r, v_gl, r1, 0);
    pd[i] = VERT ? 0.f : buf_ld(xd_r, v_x, rf);
    pu[i] = VERT ? 0.f : buf_ld(xu_r, v_x, rf);
    pw[i] = buf_ld(edge_r, v_e, qp);
    if (EXACT) {
#pragma unroll
      for (int l = 0; l < CM; ++l) pj[EXACT ? i : 0][l] = buf_ld<BPL_SEMP_AUX>(mat_r, v_mm[EXACT ? l : 0], qm);
    }
  };
  auto load_step = [&](int s, int i) {
    const uint32_t pr = pos(s + 1), pq = pos(s);
    load_at(pr * st_f, pr * st_b, pr * pstride, pq * st_p, pq * st_m,
            pr * st_f + (uds_dn && pr == 0u ? HWCu * 4u : 0u), i);
  };
  if (nsteps > 0) {
#pragma unroll
    for (int i = 0; i < PF; ++i) {
      const int s = nsteps - 1 - i;
      load_step(s >= 0 ? s : 0, i);
      __builtin_amdgcn_sched_barrier(0);  // keep the ring slots in issue order
    }
  }
  float acc = 0.f;
  int prev = SEG * RS + c;  // the routing cell written one step earlier (the dummy row at first)
  // sender offsets of the current adjoint step
  uint32_t o_f = pos(nsteps - 1) * st_f, o_p = pos(nsteps - 1) * st_p, o_m = pos(nsteps - 1) * st_m;
  auto step = [&](int i) {
    const int A = pa[i];
    const int gl = pl[i];
    const bool tr = A == kFromTrunc;
    const int r = tr ? gl : A;
    // normalisation Jacobian: g = t - softmax(msg) * sum(t), t = in_grad + chain
    const float tt = pg[i] + acc;
    const float ts = seg_reduce<SEG>(lab ? tt : 0.f, OpSum());
    float gk = tt - fast_exp(pm[i]) * ts;
    if (!VERT) gk += pd[i] + pu[i];
    asm volatile("" : "+v"(gk));
    gk = lab ? gk : 0.f;
    // route: clear last step's cell, write [r][c]; sender c sums row c in label order
    int cell = __mul24(r, RS) + c;
    asm volatile("" : "+v"(cell));
    cell = lab ? cell : SEG * RS + c;
    lds_r[seg][prev] = 0.f;
    lds_r[seg][cell] = gk;
    prev = cell;
    wave_lds_sync();
    float row[CM];
#pragma unroll
    for (int l = 0; l < CM; ++l) row[l] = lds_r[seg][c * RS + l];
    wave_lds_sync();  // the row is read before the next step's writes
    float gh = 0.f;
#pragma unroll
    for (int l = 0; l < CM; ++l) gh += row[l];
    acc = gh;
    buf_st(xo_r, v_xs, o_f, gh);
    const float w = pw[i];
    const bool to_tr = tr && EXACT;
    // edge gradient
    float ge = 0.f;
    if (EXACT) {
      // the truncation entry M[0][C-1] is the row-0 slot of the segment's lane C-1
      const int j0 = __float_as_int(pj[EXACT ? i : 0][0]);
      const float jt0 = __int_as_float(__builtin_amdgcn_readlane(j0, C - 1));
      const float jt1 = __int_as_float(__builtin_amdgcn_readlane(j0, SEG + C - 1));
      float mv = 0.f;
#pragma unroll
      for (int l = 0; l < CM; ++l) mv = l == A ? pj[EXACT ? i : 0][l] : mv;
      mv = to_tr ? (seg ? jt1 : jt0) : mv;
      ge = seg_reduce<SEG>(lab ? mv * gk : 0.f, OpSum());
    }
    buf_st(ge_r, v_ge, o_p, ge);
    // matrix gradient column c of the sender
    const float val = to_tr ? 0.f : w * gk;
    float ex = 0.f;
    if (EXACT) {
      const float fsum = seg_reduce<SEG>(lab && tr ? gk : 0.f, OpSum());
      ex = w * fsum;
    }
    const bool own_ex = EXACT && c == C - 1;
    if (IMGS) {
      float v0, v1, x0, x1;
      int r0, r1;
      swap_halves(val, v0, v1);
      swap_halves(ex, x0, x1);
      swap_halves(r, r0, r1);
#pragma unroll
      for (int jr = 0; jr < NR; ++jr) {
        const int rr = rlo + jr;
        float e0 = rr == r0 ? v0 : 0.f, e1 = rr == r1 ? v1 : 0.f;
        if (own_ex && rr == 0) {
          e0 += x0;
          e1 += x1;
        }
        buf_st<BPL_SEMP_AUX>(gm_r, v_gm[jr], o_m, e0 + e1);
      }
    } else {
#pragma unroll
      for (int rr = 0; rr < NR; ++rr) {
        float e = rr == r ? val : 0.f;
        if (own_ex && rr == 0) e += ex;
        buf_st<BPL_SEMP_AUX>(gm_r, v_gm[rr], o_m, e);
      }
    }
    o_f -= d_f;
    o_p -= d_p;
    o_m -= d_m;
  };
  // main loop: the ring refills stay within the scanline (s - PF >= 0)
  int t0 = 0;
  {
    const int s = nsteps - 1 - PF;  // the first refill's sender position
    const uint32_t pr = pos(s >= 0 ? s + 1 : 0), pq = pos(s >= 0 ? s : 0);
    uint32_t l_f = pr * st_f, l_b = pr * st_b, l_1 = pr * pstride, l_p = pq * st_p, l_m = pq * st_m;
    for (; t0 + 2 * PF <= nsteps; t0 += PF) {
#pragma unroll
      for (int i = 0; i < PF; ++i) {
        step(i);
        // refill ring slot i (its old contents are dead now: no register copies, no drain)
        load_at(l_f, l_b, l_1, l_p, l_m, l_f, i);
        l_f -= d_f;
        l_b -= d_b;
        l_1 -= d_1;
        l_p -= d_p;
        l_m -= d_m;
      }
    }
  }
  // the last groups: refills clamped to the first sender, steps past the end skipped
  for (; t0 < nsteps; t0 += PF) {
#pragma unroll
    for (int i = 0; i < PF; ++i) {
      const int t = t0 + i;
      if (t < nsteps) {
        step(i);
        const int s = nsteps - 1 - t;
        load_step(s - PF >= 0 ? s - PF : 0, i);
      }
    }
  }
}

template <int CM, int PF>
void launch_semp(const BwdArgs &a, bool vert, hipStream_t st) {
  const int L = vert ? a.W : a.H;
  const bool imgs = a.N == 2;
  const bool exact = a.mode == BPL_GRAD_EXACT;
  const dim3 grid((unsigned)(imgs ? 2 * L : a.N * 2 * ((L + 1) / 2)));
#define BPL_SEMP_LAUNCH(IM, EX)                                                                        \
  do {                                                                                                 \
    if (vert) hipLaunchKernelGGL((lbp_bwd_semp<CM, true, PF, IM, EX>), grid, dim3(kWave), 0, st, a);  \
    else hipLaunchKernelGGL((lbp_bwd_semp<CM, false, PF, IM, EX>), grid, dim3(kWave), 0, st, a);      \
  } while (0)
  if (imgs) {
    if (exact) BPL_SEMP_LAUNCH(true, true); else BPL_SEMP_LAUNCH(true, false);
  } else {
    if (exact) BPL_SEMP_LAUNCH(false, true); else BPL_SEMP_LAUNCH(false, false);
  }
#undef BPL_SEMP_LAUNCH
}

// the per-pixel kernels cover C <= 32 for batches of one or two images
bool semp_ok(const BwdArgs &a) {
  const int64_t hwc = (int64_t)a.H * a.W * a.C;
  return a.C <= 32 && a.N <= 2 && (int64_t)a.N * 4 * hwc * 4 < (1LL << 31) && hwc * a.C * 4 < (1LL << 31) &&
         !generic_only();
}

// y-boundary terms of the input-message gradient after the paired horizontal phase (which
// wrote X_R + X_L into the U and D slots): U slot of row 0 = (in_grad_U + X_D) + (X_R + X_L),
// D slot of row H - 1 = (in_grad_D + X_U) + (X_R + X_L), the combine kernel's expressions.
__global__ void lbp_bwd_pair_rows(const float *__restrict__ in_grad, const float *__restrict__ XD,
                                  const float *__restrict__ XU, float *__restrict__ grad_msg, int N, int H, int W,
                                  int C) {
  const int64_t WC = (int64_t)W * C, HWC = (int64_t)H * WC;
  const int64_t total = (int64_t)N * 2 * WC;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t n = i / (2 * WC), r = i - n * 2 * WC;
    const int side = (int)(r / WC);
    const int64_t pix = (side ? (int64_t)(H - 1) * WC : 0) + (r - side * WC);
    const int slot = side ? DOWN : UP;
    const float xv = side ? XU[n * HWC + pix] : XD[n * HWC + pix];
    float *g = grad_msg + (n * 4 + slot) * HWC + pix;
    *g = (in_grad[(n * 4 + slot) * HWC + pix] + xv) + *g;
  }
}

// grad_cost and the input-message gradient from the four sender h-gradients.
__global__ void lbp_bwd_combine(const float *__restrict__ XL, const float *__restrict__ XR,
                                const float *__restrict__ XU, const float *__restrict__ XD,
                                const float *__restrict__ in_grad, float *__restrict__ grad_cost,
                                float *grad_msg, int N, int H, int W, int C, int exact,
                                int osp, int copy_msgs, int accum_cost, int uds_out) {
  const int64_t total = (int64_t)N * H * W * C;
  const int64_t HWC = (int64_t)H * W * C;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const float xl = XL[i], xr = XR[i], xu = XU[i], xd = XD[i];
    const float gc = ((xd + xu) + xr) + xl;
    grad_cost[i] = accum_cost ? grad_cost[i] + gc : gc;
    if (!grad_msg) continue;  // the input-message gradient is not wanted
    const int64_t n = i / HWC;
    const int64_t r = i - n * HWC;            // (y, x, c) inside the image
    const int64_t base = n * 4 * HWC + r;     // slot 0 of this element
    if (!exact) {  // the reference's grad_messages = in_grad; nothing to copy when they alias
      if (copy_msgs) {
        grad_msg[base] = in_grad[base];
        grad_msg[base + HWC] = in_grad[base + HWC];
        grad_msg[base + 2 * HWC] = in_grad[base + 2 * HWC];
        grad_msg[base + 3 * HWC] = in_grad[base + 3 * HWC];
      }
      continue;
    }
    const int64_t yx = r / C;
    const int y = (int)(yx / W), x = (int)(yx - (int64_t)y * W);
    const float hr = xr + xl;   // read of U/D inputs by the horizontal passes
    const float sv = xd + xu;   // read of L/R boundary inputs by the vertical passes
    // osp: the LEFT/RIGHT slots (zero away from the row ends) are written at the row ends only
    if (!osp || x == 0) grad_msg[base + LEFT * HWC] = x == 0 ? (in_grad[base + LEFT * HWC] + xr) + sv : 0.f;
    if (!osp || x == W - 1) grad_msg[base + RIGHT * HWC] = x == W - 1 ? (in_grad[base + RIGHT * HWC] + xl) + sv : 0.f;
    grad_msg[base + UP * HWC] = y == 0 ? (in_grad[base + UP * HWC] + xd) + hr : hr;
    if (!uds_out || y == 0 || y == H - 1)  // BPL_GRAD_OUT_UD_SHARED: the D slot at rows 0, H - 1 only
      grad_msg[base + DOWN * HWC] = y == H - 1 ? (in_grad[base + DOWN * HWC] + xu) + hr : hr;
  }
}

// dst += src over n elements (BPL_GRAD_ACCUMULATE: this call's grad_jump and grad_edge)
__global__ void lbp_bwd_accumulate(float *__restrict__ dst, const float *__restrict__ src, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = dst[i] + src[i];
}

// Sum of the per-(wave, segment) C x C partials of each direction, in two deterministic stages.
// The partials are laid out in launch order: phase V then phase H; inside a phase, block b
// handles (group g = b / wpg, with direction bit g & 1), segments contiguous.  The list of a
// direction (its groups n = 0..N-1, their blocks and segments in order) is cut into
// kRedChunks chunks; stage 1 sums each chunk with 4 lanes per matrix entry (lane j takes the
// chunk's entries j, j + 4, ...; the 4 sums are combined as (s0 + s1) + (s2 + s3)); stage 2 adds
// the chunk sums in chunk order.
constexpr int kRedChunks = 16;
__global__ __launch_bounds__(256) void lbp_bwd_reduce_partials(const float *__restrict__ partial,
                                                               float *__restrict__ chunk_sums, int C,
                                                               int64_t v_base, int64_t h_base, int N, int wpg_v,
                                                               int wpg_h, int spw) {
  __shared__ float part[4][64];
  const int d = blockIdx.y, k = blockIdx.z;
  const bool vert = d == UP || d == DOWN;
  const int dsel = vert ? (d == UP) : (d == LEFT);
  const int wpg = vert ? wpg_v : wpg_h;
  const int64_t base = vert ? v_base : h_base;
  const int CC = C * C;
  const int e = blockIdx.x * 64 + (threadIdx.x & 63);
  const int j = threadIdx.x >> 6;
  const int per_n = wpg * spw;             // entries of one group in the list
  const int T = N * per_n;                 // entries of the direction
  const int i0 = (int)((int64_t)T * k / kRedChunks), i1 = (int)((int64_t)T * (k + 1) / kRedChunks);
  float sum = 0.f;
  if (e < CC) {
    for (int i = i0 + j; i < i1; i += 4) {
      const int nn = i / per_n, r = i - nn * per_n;  // group n, then (block, segment) inside it
      const int64_t slot = (int64_t)(nn * 2 + dsel) * per_n + r;
      sum += partial[base + slot * CC + e];
    }
  }
  part[j][threadIdx.x & 63] = sum;
  __syncthreads();
  if (j == 0 && e < CC) {
    const int l = threadIdx.x & 63;
    chunk_sums[((int64_t)d * kRedChunks + k) * CC + e] = (part[0][l] + part[1][l]) + (part[2][l] + part[3][l]);
  }
}
__global__ void lbp_bwd_reduce_chunks(const float *__restrict__ chunk_sums, float *__restrict__ out, int C) {
  const int CC = C * C;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 4 * CC) return;
  const int d = i / CC, e = i - d * CC;
  float s = 0.f;
  for (int k = 0; k < kRedChunks; ++k) s += chunk_sums[((int64_t)d * kRedChunks + k) * CC + e];
  out[i] = s;
}

#ifndef BPL_BWD_PAIR_LPL
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

// the fast step needs the window half-width (argmins within 3 labels), a stereo packing's
// jump length and 32-bit buffer offsets into the (direction) planes of the pairwise gradient
// and 32-bit byte offsets into every [H,W,C] plane (its row streams are buffer-addressed)
bool stereo_fast_ok(const BwdArgs &a) {
  return a.delta >= 0 && a.delta <= 3 && !a.accum && (a.K == 6 || a.K == 8) &&
         (int64_t)a.K * a.H * a.W * 4 < (1LL << 31) && (int64_t)a.H * a.W * a.C * 4 < (1LL << 31);
}
#ifndef BPL_BWD_PF_V
#define BPL_BWD_PF_V 4
#endif
template <int LPL, bool FULL>
void launch_stereo_w64(const BwdArgs &a, bool vert, bool pair, int ngroups, hipStream_t st) {
  // prefetch depth of the vertical phase (the horizontal kernels: launch_bwd_pair,
  // launch_bwd_h_unpaired)
  constexpr int PF = BPL_BWD_PF_V;
  const bool fast = stereo_fast_ok(a);
  auto go = [&](auto fk) {
    constexpr int FK = decltype(fk)::value;
    if (vert) {
      const dim3 grid((unsigned)(ngroups * 2 * a.W));
      hipLaunchKernelGGL((lbp_bwd_stereo_w64<LPL, true, FULL, FK, false, PF>), grid, dim3(kWave), 0, st, a);
    } else if (pair) {
      // pairing needs the fast step (stereo_pair_ok), so FK > 0
      launch_bwd_pair<LPL>(a, FULL, FK, ngroups, st);
    } else {
      launch_bwd_h_unpaired<LPL>(a, FULL, FK, ngroups, st);
    }
  };
  if (fast && a.K == 6) go(std::integral_constant<int, 6>());
  else if (fast && a.K == 8) go(std::integral_constant<int, 8>());
  else go(std::integral_constant<int, 0>());
}
template <int LPL>
void launch_stereo_w64_any(const BwdArgs &a, bool vert, bool pair, int ngroups, hipStream_t st) {
  // full rows of 4 labels per lane would be C = 256, above the 255-label limit
  if (LPL < 4 && a.C == kWave * LPL) launch_stereo_w64<LPL, LPL < 4>(a, vert, pair, ngroups, st);
  else launch_stereo_w64<LPL, false>(a, vert, pair, ngroups, st);
}
// the 64-lane stereo kernel covers 16 < C <= 256 and up to 8 pairwise bins
bool stereo_w64_ok(const BwdArgs &a) { return a.C > 16 && a.K <= 8; }
// exact-mode stereo backward whose horizontal phase pairs the two passes of a row and writes
// grad_cost and the input-message gradient itself (no combine pass)
bool stereo_pair_ok(const BwdArgs &a) {
  return stereo_w64_ok(a) && !generic_only() && a.mode == BPL_GRAD_EXACT && stereo_fast_ok(a) && a.W >= 2;
}
void launch_stereo(const BwdArgs &a, bool vert, bool pair, int ngroups, hipStream_t st) {
  const int C = a.C;
  if (!stereo_w64_ok(a) || generic_only()) launch_kind<BPL_LBP_STEREO>(a, vert, ngroups, st);
  else if (C <= 64) launch_stereo_w64_any<1>(a, vert, pair, ngroups, st);
  else if (C <= 128) launch_stereo_w64_any<2>(a, vert, pair, ngroups, st);
  else if (C <= 192) launch_stereo_w64_any<3>(a, vert, pair, ngroups, st);
  else launch_stereo_w64_any<4>(a, vert, pair, ngroups, st);
}

int seg_width(int C) { return C <= 16 ? 16 : C <= 32 ? 32 : 64; }
// the semantic-global kernel for 16 < C <= 32: its partials use the 32-lane segments of the
// generic layout
bool semg_ok(const BwdArgs &a) {
  return a.C > 16 && a.C <= 32 && (int64_t)a.H * a.W * a.C * 4 < (1LL << 31) && !generic_only();
}

// BPL_GRAD_ACCUMULATE: the kernels write this call's grad_jump and grad_edge into workspace
// copies, which lbp_bwd_accumulate then adds to the caller's tensors (stereo jumps of up to
// 15 entries for each of N images; semantic global: one 4 x C x C matrix set)
constexpr int kMaxJumpK = 15;
int64_t acc_jump_floats(int kind, int64_t N, int64_t H, int64_t W, int64_t C) {
  if (kind == BPL_LBP_STEREO) return N * 4 * kMaxJumpK * H * W;
  if (kind == BPL_LBP_SEMANTIC_GLOBAL) return 4 * C * C;
  return 0;
}
int64_t acc_floats(int kind, int64_t N, int64_t H, int64_t W, int64_t C) {
  return kind == BPL_LBP_SEMANTIC_PIXEL ? 0 : acc_jump_floats(kind, N, H, W, C) + N * 4 * H * W;
}
void ws_layout(int kind, int64_t N, int64_t H, int64_t W, int64_t C, int64_t *planes, int64_t *part_v,
               int64_t *part_h, int64_t *total) {
  const int64_t plane = N * H * W * C;
  *planes = 4 * plane;
  int64_t pv = 0, ph = 0;
  if (kind == BPL_LBP_SEMANTIC_GLOBAL) {
    const int spw = kWave / seg_width((int)C);
    const int64_t wv = N * 2 * ((W + spw - 1) / spw), wh = N * 2 * ((H + spw - 1) / spw);
    pv = wv * spw * C * C;
    ph = wh * spw * C * C + 4 * kRedChunks * C * C;  // then the reduction's chunk sums
  }
  *part_v = pv;
  *part_h = ph;
  *total = (*planes + pv + ph + acc_floats(kind, N, H, W, C)) * (int64_t)sizeof(float);
}

}  // namespace
}  // namespace bpl

using namespace bpl;

extern "C" int64_t bpl_lbp_bwd_workspace_size(int kind, int64_t N, int64_t H, int64_t W, int64_t C) {
  int64_t planes, pv, ph, total;
  ws_layout(kind, N, H, W, C, &planes, &pv, &ph, &total);
  return total;
}

namespace {

int lbp_bwd_impl(int kind, int mode, const float *jump, int64_t jump_k, int64_t jump_n,
                 const float *edge, const float *in_grad, const float *messages,
                 const uint8_t *argmin, const uint8_t *gargmin,
                 float *grad_cost, float *grad_jump, float *grad_edge, float *grad_msg,
                 void *workspace, int64_t workspace_bytes,
                 int64_t N, int64_t H, int64_t W, int64_t C, int delta, void *stream) {
  BPL_REQUIRE(kind >= 0 && kind <= 2, BPL_E_ARG, "bpl_lbp_bwd: unknown kind %d", kind);
  const int hflags = mode & (BPL_GRAD_IN_H_ROW_ENDS | BPL_GRAD_OUT_H_ROW_ENDS | BPL_GRAD_IN_UD_SHARED |
                            BPL_GRAD_OUT_UD_SHARED);
  const bool accumulate = (mode & BPL_GRAD_ACCUMULATE) != 0;
  mode &= ~(BPL_GRAD_IN_H_ROW_ENDS | BPL_GRAD_OUT_H_ROW_ENDS | BPL_GRAD_IN_UD_SHARED | BPL_GRAD_OUT_UD_SHARED |
            BPL_GRAD_ACCUMULATE);
  BPL_REQUIRE(!accumulate || kind != BPL_LBP_SEMANTIC_PIXEL, BPL_E_ARG,
              "bpl_lbp_bwd: BPL_GRAD_ACCUMULATE applies to the stereo and semantic-global kinds");
  BPL_REQUIRE(mode == BPL_GRAD_REFERENCE || mode == BPL_GRAD_EXACT, BPL_E_ARG, "bpl_lbp_bwd: unknown mode %d", mode);
  BPL_REQUIRE(hflags == 0 || mode == BPL_GRAD_EXACT, BPL_E_ARG,
              "bpl_lbp_bwd: the row-end and shared-slot flags apply to the exact mode only");
  BPL_REQUIRE(N >= 1 && H >= 1 && W >= 1 && C >= 1, BPL_E_ARG, "bpl_lbp_bwd: sizes must be positive");
  BPL_REQUIRE(C <= 255, BPL_E_LIMIT, "bpl_lbp_bwd: C=%lld labels exceeds the supported 255", (long long)C);
  BPL_REQUIRE(N * 2 * (H > W ? H : W) < (1LL << 31) && H < (1 << 30) && W < (1 << 30), BPL_E_LIMIT,
              "bpl_lbp_bwd: grid too large");
  if (grad_msg) {
    // grad_msg and in_grad are [N,4,H,W,C] each: they must be disjoint, or (reference mode
    // only, where nothing is written to grad_msg) the same tensor; a partial overlap would
    // race with kernels that read in_grad while others write grad_msg
    const uintptr_t g0 = (uintptr_t)grad_msg, i0 = (uintptr_t)in_grad;
    const uintptr_t bytes = (uintptr_t)(N * 4 * H * W * C) * sizeof(float);
    const bool same = g0 == i0, disjoint = g0 + bytes <= i0 || i0 + bytes <= g0;
    BPL_REQUIRE(disjoint || (same && mode == BPL_GRAD_REFERENCE), BPL_E_ARG,
                same ? "bpl_lbp_bwd: grad_msg may alias in_grad in the reference mode only"
                     : "bpl_lbp_bwd: grad_msg partly overlaps in_grad");
  }
  BPL_REQUIRE(edge && in_grad && messages && argmin && gargmin && grad_cost && grad_jump && grad_edge,
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
  a.delta = delta;
  a.igz = (hflags & BPL_GRAD_IN_H_ROW_ENDS) != 0;
  a.osp = (hflags & BPL_GRAD_OUT_H_ROW_ENDS) != 0;
  a.uds_in = (hflags & BPL_GRAD_IN_UD_SHARED) != 0;
  a.uds_out = (hflags & BPL_GRAD_OUT_UD_SHARED) != 0;
  a.grad_cost = grad_cost; a.grad_msg = grad_msg;
  a.accum_cost = accumulate ? 1 : 0;
  // accumulating: this call's pairwise and edge gradients go to workspace copies first
  float *acc_jump = ws + planes + pv + ph;
  float *acc_edge = acc_jump + acc_jump_floats(kind, N, H, W, C);
  const int64_t jump_elems = kind == BPL_LBP_STEREO ? jump_n * 4 * jump_k * H * W : 4 * C * C;
  const int64_t edge_elems = N * 4 * H * W;
  if (accumulate) {
    a.grad_jump = acc_jump;
    a.grad_edge = acc_edge;
  }
  float *out_jump = a.grad_jump;
  hipStream_t st = as_stream(stream);
  // parameter gradients shared across the batch accumulate over n in launch order
  const bool semp = kind == BPL_LBP_SEMANTIC_PIXEL && semp_ok(a);
  const bool per_n = (kind == BPL_LBP_STEREO && a.Nj == 1 && N > 1) || (kind == BPL_LBP_SEMANTIC_PIXEL && N > 1 && !semp);
  a.accum = 0;
  const bool pair = kind == BPL_LBP_STEREO && !per_n && stereo_pair_ok(a);
  for (int phase = 0; phase < 2; ++phase) {
    const bool vert = phase == 0;
    if (kind == BPL_LBP_SEMANTIC_GLOBAL) a.partial = ws + planes + (vert ? 0 : pv);
    const int nb = per_n ? (int)N : 1;
    KernelTimer kt(vert ? BPL_KERNEL_LBP_BWD_V : BPL_KERNEL_LBP_BWD_H, st);
    for (int b = 0; b < nb; ++b) {
      a.n_base = per_n ? b : 0;
      a.accum = per_n && b > 0;
      const int ng = per_n ? 1 : (int)N;
      switch (kind) {
        case BPL_LBP_STEREO: launch_stereo(a, vert, pair, ng, st); break;
        case BPL_LBP_SEMANTIC_GLOBAL:
          if (semg_ok(a) && C <= 20) launch_semg<20>(a, vert, ng, st);
          else if (semg_ok(a)) launch_semg<32>(a, vert, ng, st);
          else launch_kind<BPL_LBP_SEMANTIC_GLOBAL>(a, vert, ng, st);
          break;
        default:
          if (semp && C <= 20) launch_semp<20, 4>(a, vert, st);
          else if (semp) launch_semp<32, 2>(a, vert, st);
          else launch_kind<BPL_LBP_SEMANTIC_PIXEL>(a, vert, ng, st);
          break;
      }
      int rc = check_hip(hipGetLastError(), "bpl_lbp_bwd phase launch");
      if (rc) return rc;
    }
  }
  if (pair && !grad_msg) {
    // the paired horizontal phase wrote grad_cost; the input-message gradient is not wanted
  } else if (pair) {
    // the paired horizontal phase wrote grad_cost and grad_msg; add the y-boundary terms
    const int64_t total_el = N * 2 * W * C;
    const int threads = 256;
    int64_t blocks = (total_el + threads - 1) / threads;
    if (blocks > 65536) blocks = 65536;
    KernelTimer kt(BPL_KERNEL_LBP_BWD_COMBINE, st);
    hipLaunchKernelGGL(lbp_bwd_pair_rows, dim3((unsigned)blocks), dim3(threads), 0, st, in_grad, a.X[DOWN], a.X[UP],
                       grad_msg, (int)N, (int)H, (int)W, (int)C);
    int rc = check_hip(hipGetLastError(), "bpl_lbp_bwd boundary-row launch");
    if (rc) return rc;
  } else {
    const int64_t total_el = plane;
    const int threads = 256;
    int64_t blocks = (total_el + threads - 1) / threads;
    if (blocks > 65536) blocks = 65536;
    KernelTimer kt(BPL_KERNEL_LBP_BWD_COMBINE, st);
    hipLaunchKernelGGL(lbp_bwd_combine, dim3((unsigned)blocks), dim3(threads), 0, st, a.X[LEFT], a.X[RIGHT],
                       a.X[UP], a.X[DOWN], in_grad, grad_cost, grad_msg, (int)N, (int)H, (int)W, (int)C,
                       mode == BPL_GRAD_EXACT ? 1 : 0, a.osp, grad_msg != in_grad ? 1 : 0, a.accum_cost, a.uds_out);
    int rc = check_hip(hipGetLastError(), "bpl_lbp_bwd combine launch");
    if (rc) return rc;
  }
  if (kind == BPL_LBP_SEMANTIC_GLOBAL) {
    const int spw = kWave / seg_width((int)C);
    const int wpg_v = (int)((W + spw - 1) / spw), wpg_h = (int)((H + spw - 1) / spw);
    KernelTimer kt(BPL_KERNEL_LBP_BWD_REDUCE, st);
    float *chunk_sums = ws + planes + pv + (ph - 4 * kRedChunks * C * C);
    hipLaunchKernelGGL(lbp_bwd_reduce_partials, dim3((unsigned)((C * C + 63) / 64), 4, kRedChunks), dim3(256), 0, st,
                       ws + planes, chunk_sums, (int)C, (int64_t)0, pv, (int)N, wpg_v, wpg_h, spw);
    int rc = check_hip(hipGetLastError(), "bpl_lbp_bwd reduce launch");
    if (rc) return rc;
    hipLaunchKernelGGL(lbp_bwd_reduce_chunks, dim3((unsigned)((4 * C * C + 255) / 256)), dim3(256), 0, st,
                       chunk_sums, out_jump, (int)C);
    rc = check_hip(hipGetLastError(), "bpl_lbp_bwd reduce launch");
    if (rc) return rc;
  }
  if (accumulate) {
    KernelTimer kt(BPL_KERNEL_LBP_BWD_COMBINE, st);
    auto add = [&](float *dst, const float *src, int64_t n) {
      int64_t blocks = (n + 255) / 256;
      if (blocks > 65536) blocks = 65536;
      hipLaunchKernelGGL(lbp_bwd_accumulate, dim3((unsigned)blocks), dim3(256), 0, st, dst, src, n);
      return check_hip(hipGetLastError(), "bpl_lbp_bwd accumulate launch");
    };
    int rc = add(grad_jump, acc_jump, jump_elems);
    if (rc) return rc;
    rc = add(grad_edge, acc_edge, edge_elems);
    if (rc) return rc;
  }
  return BPL_OK;
}

// Scratch memory of bpl_lbp_bwd_noalloc: one array per GPU, in the code object, so the entry
// point needs no allocation.  In host code `scratch` names a host-side stand-in: its device
// address comes from hipGetSymbolAddress, made once per device with that device current.
constexpr int64_t kScratchBytes = BPL_LBP_BWD_NOALLOC_MAX_BYTES;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

int scratch_of_current_device(void **out) {
  int dev = -1;
  int rc = check_hip(hipGetDevice(&dev), "bpl_lbp_bwd_noalloc: hipGetDevice");
  if (rc) return rc;
  BPL_REQUIRE(dev >= 0 && dev < kMaxDevices, BPL_E_LIMIT, "bpl_lbp_bwd_noalloc: device %d beyond %d", dev,
              kMaxDevices);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    rc = check_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "bpl_lbp_bwd_noalloc: hipGetSymbolAddress");
    if (rc) return rc;
    BPL_REQUIRE(p != nullptr, BPL_E_HIP, "bpl_lbp_bwd_noalloc: no device address for the scratch array");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return BPL_OK;
}

}  // namespace

extern "C" int bpl_lbp_bwd(int kind, int mode, const float *jump, int64_t jump_k, int64_t jump_n,
                           const float *edge, const float *in_grad, const float *messages,
                           const uint8_t *argmin, const uint8_t *gargmin,
                           float *grad_cost, float *grad_jump, float *grad_edge, float *grad_msg,
                           void *workspace, int64_t workspace_bytes,
                           int64_t N, int64_t H, int64_t W, int64_t C, int delta, void *stream) {
  return lbp_bwd_impl(kind, mode, jump, jump_k, jump_n, edge, in_grad, messages, argmin, gargmin, grad_cost,
                      grad_jump, grad_edge, grad_msg, workspace, workspace_bytes, N, H, W, C, delta, stream);
}

extern "C" int bpl_lbp_bwd_noalloc(int kind, int mode, const float *jump, int64_t jump_k, int64_t jump_n,
                                   const float *edge, const float *in_grad, const float *messages,
                                   const uint8_t *argmin, const uint8_t *gargmin,
                                   float *grad_cost, float *grad_jump, float *grad_edge, float *grad_msg,
                                   int64_t N, int64_t H, int64_t W, int64_t C, int delta, void *stream) {
  BPL_REQUIRE(kind >= 0 && kind <= 2, BPL_E_ARG, "bpl_lbp_bwd_noalloc: unknown kind %d", kind);
  BPL_REQUIRE(N >= 1 && H >= 1 && W >= 1 && C >= 1 && C <= 255, BPL_E_ARG,
              "bpl_lbp_bwd_noalloc: sizes must be positive and C <= 255");
  int64_t planes, pv, ph, total;
  ws_layout(kind, N, H, W, C, &planes, &pv, &ph, &total);
  BPL_REQUIRE(total <= kScratchBytes, BPL_E_LIMIT,
              "bpl_lbp_bwd_noalloc: these sizes need %lld bytes of workspace, above the %lld bytes of the "
              "built-in scratch array (use bpl_lbp_bwd with a caller-owned workspace)",
              (long long)total, (long long)kScratchBytes);
  void *ws = nullptr;
  int rc = scratch_of_current_device(&ws);
  if (rc) return rc;
  return lbp_bwd_impl(kind, mode, jump, jump_k, jump_n, edge, in_grad, messages, argmin, gargmin, grad_cost,
                      grad_jump, grad_edge, grad_msg, ws, kScratchBytes, N, H, W, C, delta, stream);
}

#else  // BPL_BWD_PAIR_LPL: the paired horizontal launches for one labels-per-lane count

}  // namespace

// Prefetch depth 8: the paired kernel has one wave per SIMD, so registers are spare; with
// default-policy row loads a shallower ring left its steps waiting on memory (384x1248x192:
// 1.205 ms per launch with 4 steps, 1.147 with 8, 1.084 with 12), with non-temporal ones 8 and
// 12 steps are within 0.5% (8 ahead) and 16 is slower.  HS carries the chained-iteration flags.
template <int LPL>
void launch_bwd_pair(const BwdArgs &a, bool full, int fk, int ngroups, hipStream_t st) {
  constexpr int PFH = 8;
  const dim3 grid((unsigned)(ngroups * ((a.H + 1) / 2)));  // two rows per workgroup
  const int hs = (a.igz ? 1 : 0) | (a.osp ? 2 : 0);
  auto go = [&](auto full_c, auto fk_c) {
    constexpr bool FULL = decltype(full_c)::value;
    constexpr int FK = decltype(fk_c)::value;
    if (hs == 0)
      hipLaunchKernelGGL((lbp_bwd_stereo_w64<LPL, false, FULL, FK, true, PFH, 0>), grid, dim3(4 * kWave), 0, st, a);
    else if (hs == 1)
      hipLaunchKernelGGL((lbp_bwd_stereo_w64<LPL, false, FULL, FK, true, PFH, 1>), grid, dim3(4 * kWave), 0, st, a);
    else if (hs == 2)
      hipLaunchKernelGGL((lbp_bwd_stereo_w64<LPL, false, FULL, FK, true, PFH, 2>), grid, dim3(4 * kWave), 0, st, a);
    else
      hipLaunchKernelGGL((lbp_bwd_stereo_w64<LPL, false, FULL, FK, true, PFH, 3>), grid, dim3(4 * kWave), 0, st, a);
  };
  using T = std::true_type;
  using F = std::false_type;
  using K6 = std::integral_constant<int, 6>;
  using K8 = std::integral_constant<int, 8>;
  // full rows of 4 labels per lane would be C = 256, above the 255-label limit
  if constexpr (LPL < 4) {
    if (full) {
      if (fk == 6) go(T(), K6());
      else go(T(), K8());
      return;
    }
  }
  if (fk == 6) go(F(), K6());
  else go(F(), K8());
}
template void launch_bwd_pair<BPL_BWD_PAIR_LPL>(const BwdArgs &, bool, int, int, hipStream_t);

// Unpaired horizontal adjoints (one wave per row and pass): the fast step with a 12-step ring
// (one wave per SIMD here too), the general step with 4.
template <int LPL>
void launch_bwd_h_unpaired(const BwdArgs &a, bool full, int fk, int ngroups, hipStream_t st) {
  const dim3 grid((unsigned)(ngroups * 2 * a.H));
  auto go = [&](auto full_c, auto fk_c) {
    constexpr bool FULL = decltype(full_c)::value;
    constexpr int FK = decltype(fk_c)::value;
    constexpr int PFU = FK > 0 ? 12 : 4;
    hipLaunchKernelGGL((lbp_bwd_stereo_w64<LPL, false, FULL, FK, false, PFU>), grid, dim3(kWave), 0, st, a);
  };
  using T = std::true_type;
  using F = std::false_type;
  auto by_k = [&](auto full_c) {
    if (fk == 6) go(full_c, std::integral_constant<int, 6>());
    else if (fk == 8) go(full_c, std::integral_constant<int, 8>());
    else go(full_c, std::integral_constant<int, 0>());
  };
  if constexpr (LPL < 4) {
    if (full) {
      by_k(T());
      return;
    }
  }
  by_k(F());
}
template void launch_bwd_h_unpaired<BPL_BWD_PAIR_LPL>(const BwdArgs &, bool, int, int, hipStream_t);

}  // namespace bpl
#endif  // BPL_BWD_PAIR_LPL
