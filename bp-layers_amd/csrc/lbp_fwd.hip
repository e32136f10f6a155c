// Min-sum LBP forward sweeps on gfx950.
//
// Semantics follow the reference kernel lbp_cuda_forward_kernel_reduction_min_sum
// (ops/lbp_stereo/src/lbp_min_sum_kernel.cu:9-213 and its semantic twins) and the host
// loop (:424-500); the design is our own:
//   * one launch per phase (horizontal: RIGHT+LEFT, vertical: DOWN+UP) instead of one
//     launch per column/row; each wave walks whole scanlines, carrying the chain message
//     in registers;
//   * labels live on lanes (SEG lanes per scanline, LPL labels per lane); small label
//     counts pack 64/SEG scanlines into one wave;
//   * reductions over labels are DPP row stages folded into the arithmetic plus the gfx950
//     row/half permlane swaps; the generic kernel reads the candidate window from a
//     per-segment LDS copy of h, the stereo kernels from neighbouring lanes (DPP shifts);
//   * the next PF steps' streamed inputs (cost and the two cross messages) are prefetched
//     into a register ring.
#include <type_traits>

#include "bpl_common.h"

// cache policy of the semantic per-pixel forward's matrix-row loads: default (non-temporal
// made both phases 8-11% slower)
#ifndef BPL_SEMP_AUX
#define BPL_SEMP_AUX 0
#endif

namespace bpl {
namespace {

struct FwdArgs {
  const float *cost, *jump, *edge, *msg_in;
  float *msg_out;
  uint8_t *argmin, *gargmin;
  float *scale;
  int N, H, W, C, K, Nj, delta, bits, mat_lds;
  int buf_ok;  // every [H,W,C] plane and the K jump planes of a direction below 2^31 bytes
  // [N,H,W,C] plane (cost + M_R) + M_L: written by the paired horizontal kernel, read by the
  // vertical P1/P2 sweep in place of the cost and both horizontal messages (null: not used)
  float *base;
  // msg_in is all zeros (the initial messages of an iteration chain): the P1/P2 sweeps read
  // them through a resource of 0 bytes, which returns zeros without memory traffic
  int in_zero;
};

__device__ __forceinline__ int64_t px(int n, int y, int x, int H, int W) {
  return ((int64_t)n * H + y) * (int64_t)W + x;
}
__device__ __forceinline__ int64_t sl_px(int n, int s, int y, int x, int H, int W) {
  return (((int64_t)n * 4 + s) * H + y) * (int64_t)W + x;
}
__device__ __forceinline__ int sat_u16(float v) { return (int)fminf(fmaxf(v, 0.f), 65535.f); }

template <int KIND, int SEG, int LPL, bool VERT, int PF>
__global__ __launch_bounds__(64) void lbp_fwd_phase(FwdArgs a) {
  constexpr int SPW = kWave / SEG;
  constexpr int CP = SEG * LPL;
  __shared__ float lds_h[2][SPW][CP];
  __shared__ float lds_p[2][SPW][16];
  extern __shared__ float lds_m[];  // semantic global: the C x C matrix of this direction

  const int H = a.H, W = a.W, C = a.C, K = a.K;
  const int L = VERT ? W : H;               // scanlines per (n, direction)
  const int len = VERT ? H : W;             // senders + 1 per scanline
  const int nsteps = len - 1;
  const int wpg = (L + SPW - 1) / SPW;
  // columns of a vertical launch go to the XCDs in contiguous ranges (shared cache lines)
  const int bi = VERT ? xcd_contiguous_block((int)blockIdx.x, (int)gridDim.x) : (int)blockIdx.x;
  const int g = bi / wpg;
  const int wl = bi - g * wpg;
  const int n = g >> 1;
  const int dsel = g & 1;
  const int dir = VERT ? (dsel ? UP : DOWN) : (dsel ? LEFT : RIGHT);
  const int rslot = dir == RIGHT ? LEFT : dir == LEFT ? RIGHT : dir == DOWN ? UP : DOWN;
  const bool fwd = (dir == RIGHT || dir == DOWN);
  const int lane = threadIdx.x;
  const int seg = lane / SEG;
  const int sl = lane - seg * SEG;
  const int line = wl * SPW + seg;
  const bool line_ok = line < L;
  const int ln = line_ok ? line : 0;
  const bool trunc = a.delta < C;
  const int delta = a.delta;
  const int nj = a.Nj > 1 ? n : 0;
  const int s1 = VERT ? RIGHT : UP;         // cross slots, in the reference's summation order
  const int s2 = VERT ? LEFT : DOWN;
  const float *cross = VERT ? a.msg_out : a.msg_in;

  if (KIND == BPL_LBP_SEMANTIC_GLOBAL && a.mat_lds) {
    const float *M = a.jump + (int64_t)dir * C * C;
    for (int i = lane; i < C * C; i += kWave) lds_m[i] = M[i];
    __syncthreads();
  }

  auto pos_of = [&](int s) { return fwd ? s : len - 1 - s; };
  auto yx_of = [&](int p, int &y, int &x) {
    if (VERT) { y = p; x = ln; } else { y = ln; x = p; }
  };

  // ---- boundary: the first sender's chain slot is never written by this pass
  float chain[LPL];
  {
    int y, x;
    yx_of(pos_of(0), y, x);
    const int64_t o = sl_px(a.N > 0 ? n : 0, rslot, y, x, H, W);
#pragma unroll
    for (int k = 0; k < LPL; ++k) {
      const int c = sl + SEG * k;
      float v = 0.f;
      if (line_ok && c < C) {
        v = a.msg_in[o * C + c];
        a.msg_out[o * C + c] = v;
        if (a.argmin) a.argmin[o * C + c] = 0;
      }
      chain[k] = v;
    }
    if (line_ok && sl == 0) {
      if (a.scale) a.scale[o] = 0.f;
      if (a.gargmin) a.gargmin[o] = 0;
    }
  }

  // ---- register prefetch ring
  float pc[PF][LPL], p1[PF][LPL], p2[PF][LPL], pp[PF];
  auto load_step = [&](int s, float *dc, float *d1, float *d2, float &dp) {
    int y, x;
    yx_of(pos_of(s), y, x);
    const int64_t cp = px(n, y, x, H, W) * C;
    const int64_t o1 = sl_px(n, s1, y, x, H, W) * C;
    const int64_t o2 = sl_px(n, s2, y, x, H, W) * C;
#pragma unroll
    for (int k = 0; k < LPL; ++k) {
      const int c = min(sl + SEG * k, C - 1);  // clamped: loads stay unconditional
      dc[k] = a.cost[cp + c];
      d1[k] = cross[o1 + c];
      d2[k] = cross[o2 + c];
    }
    const int64_t eo = sl_px(n, dir, y, x, H, W);
    const int j = (KIND == BPL_LBP_STEREO && sl < K) ? sl : -1;
    dp = j >= 0 ? a.jump[(((int64_t)nj * 4 + dir) * K + j) * (int64_t)H * W + (int64_t)y * W + x] : a.edge[eo];
  };

#pragma unroll
  for (int i = 0; i < PF; ++i) {
    load_step(i < nsteps ? i : (nsteps > 0 ? nsteps - 1 : 0), pc[i], p1[i], p2[i], pp[i]);
    // keep the slots in issue order: the loop's wait counts assume slot 0 is the oldest
    __builtin_amdgcn_sched_barrier(0);
  }

  for (int s0 = 0; s0 < nsteps; s0 += PF) {
#pragma unroll
    for (int i = 0; i < PF; ++i) {
      const int s = s0 + i;
      if (s < nsteps) {
        const float pv = pp[i];
        const int buf = s & 1;
        int y, x;
        yx_of(pos_of(s), y, x);

        float h[LPL];
#pragma unroll
        for (int k = 0; k < LPL; ++k) {
          h[k] = VERT ? ((pc[i][k] + p1[i][k]) + p2[i][k]) + chain[k] : ((pc[i][k] + chain[k]) + p1[i][k]) + p2[i][k];
          const int c = sl + SEG * k;
          if (c < C) lds_h[buf][seg][c] = h[k];
        }
        if (sl < 16) lds_p[buf][seg][sl] = pv;
        __syncthreads();

        const float w = lds_p[buf][seg][KIND == BPL_LBP_STEREO ? K : 0];
        float msg[LPL];
        int arg[LPL];
        int lstar = 0;
        if (trunc) {
          float bv = __builtin_inff();
          int bk = -1;
#pragma unroll
          for (int k = 0; k < LPL; ++k) {
            const int c = sl + SEG * k;
            if (c < C) argmin_step(bv, bk, h[k], tie_key(c, a.bits));
          }
          seg_argmin<SEG>(bv, bk);
          lstar = tie_key(bk, a.bits);  // bit reversal is an involution
          float JK;
          if (KIND == BPL_LBP_STEREO) JK = lds_p[buf][seg][K - 1];
          else if (KIND == BPL_LBP_SEMANTIC_GLOBAL) JK = a.mat_lds ? lds_m[C - 1] : a.jump[(int64_t)dir * C * C + C - 1];
          else JK = a.jump[px(dir, y, x, H, W) * C * C + C - 1];
          const float gv = fmaf(JK, w, bv);
#pragma unroll
          for (int k = 0; k < LPL; ++k) { msg[k] = gv; arg[k] = kFromTrunc; }
          if (KIND == BPL_LBP_STEREO) {
            const float o = lds_p[buf][seg][0];
            int k_lo = -delta, k_hi = delta;
            if (o != 0.f) { k_lo = (int)floorf(o - 0.5f) - delta; k_hi = (int)ceilf(o + 0.5f) + delta; }
            int st[LPL], sp[LPL];
#pragma unroll
            for (int k = 0; k < LPL; ++k) {
              const int c = sl + SEG * k;
              st[k] = sat_u16(roundf((float)(c - delta + 1) + o - 1.0f));
              sp[k] = min(sat_u16(roundf((float)(c + delta) + o + 1.0f)), C);
            }
            for (int kk = k_lo; kk <= k_hi; ++kk) {
              const float fk = lds_p[buf][seg][min(max(K / 2 - kk, 1), K - 1)];
#pragma unroll
              for (int k = 0; k < LPL; ++k) {
                const int c = sl + SEG * k;
                const int l = c + kk;
                if (c < C && l >= st[k] && l < sp[k]) {
                  const float nv = fmaf(fk, w, lds_h[buf][seg][l]);
                  if (nv <= msg[k]) { msg[k] = nv; arg[k] = l; }
                }
              }
            }
          } else {
            for (int kk = 1 - delta; kk <= delta - 1; ++kk) {
#pragma unroll
              for (int k = 0; k < LPL; ++k) {
                const int c = sl + SEG * k;
                const int l = c + kk;
                if (c < C && l >= 0 && l < C) {
                  float fv;
                  if (KIND == BPL_LBP_SEMANTIC_GLOBAL)
                    fv = a.mat_lds ? lds_m[l * C + c] : a.jump[((int64_t)dir * C + l) * C + c];
                  else
                    fv = a.jump[(px(dir, y, x, H, W) * C + l) * C + c];
                  const float nv = fmaf(fv, w, lds_h[buf][seg][l]);
                  if (nv <= msg[k]) { msg[k] = nv; arg[k] = l; }
                }
              }
            }
          }
        } else {
#pragma unroll
          for (int k = 0; k < LPL; ++k) { msg[k] = 1e15f; arg[k] = 0; }
          const float *Mpix = (KIND == BPL_LBP_SEMANTIC_PIXEL) ? a.jump + px(dir, y, x, H, W) * C * C : nullptr;
          for (int l = 0; l < C; ++l) {
            const float hv = lds_h[buf][seg][l];
#pragma unroll
            for (int k = 0; k < LPL; ++k) {
              const int c = sl + SEG * k;
              if (c < C) {
                float fv;
                if (KIND == BPL_LBP_STEREO) fv = lds_p[buf][seg][min(max(K / 2 + c - l, 1), K - 1)];
                else if (KIND == BPL_LBP_SEMANTIC_GLOBAL)
                  fv = a.mat_lds ? lds_m[l * C + c] : a.jump[((int64_t)dir * C + l) * C + c];
                else fv = Mpix[l * C + c];
                const float nv = fmaf(fv, w, hv);
                if (nv <= msg[k]) { msg[k] = nv; arg[k] = l; }
              }
            }
          }
        }

        // ---- normalisation: m - log sum_c exp(m), max-shifted
        float mx = -__builtin_inff();
#pragma unroll
        for (int k = 0; k < LPL; ++k)
          if (sl + SEG * k < C) mx = fmaxf(mx, msg[k]);
        mx = seg_reduce<SEG>(mx, OpMax());
        float se = 0.f;
#pragma unroll
        for (int k = 0; k < LPL; ++k)
          if (sl + SEG * k < C) se += fast_exp(msg[k] - mx);
        se = seg_reduce<SEG>(se, OpSum());
        const float lse = mx + fast_log(se);

        int yq = y, xq = x;
        if (VERT) yq += fwd ? 1 : -1; else xq += fwd ? 1 : -1;
        const int64_t oq = sl_px(n, rslot, yq, xq, H, W);
#pragma unroll
        for (int k = 0; k < LPL; ++k) {
          const int c = sl + SEG * k;
          const float m = msg[k] - lse;
          chain[k] = m;
          if (line_ok && c < C) {
            a.msg_out[oq * C + c] = m;
            if (a.argmin) a.argmin[oq * C + c] = (uint8_t)arg[k];
          }
        }
        if (line_ok && sl == 0) {
          if (a.scale) a.scale[oq] = fast_exp(lse);
          if (a.gargmin) a.gargmin[oq] = (uint8_t)(trunc ? lstar : 0);
        }
      }
      // refill ring slot i (its old contents are dead now: no register copies, no drain)
      load_step(s + PF < nsteps ? s + PF : nsteps - 1, pc[i], p1[i], p2[i], pp[i]);
    }
  }
}

// ---------------------------------------------------------------------------------------
// Semantic global pairwise, full min-plus branch (delta >= C), C <= CM (20 or 32): two
// scanlines per wave (32 lanes each, label c = lane & 31).  Each lane keeps the column
// M[dir][.][c] of its label in registers (CM values, 0 past C); per step the segment's h row
// goes through LDS (+inf past C) and comes back as wave-uniform vector reads, the CM
// candidates h[l] + w M[l][c] form a min3 tree, and the winner is the last l equal to the
// minimum (the reference's `<=` scan from msg = 1e15 in label order).  Outputs as in
// lbp_fwd_phase: msg - lse (max-shifted log-sum-exp), scale = exp(lse), label of the minimum 0.
//
// The chain runs in a shifted frame: the next sender's h takes u = msg - min_l h[l] in place of
// the normalised msg - lse.  A min-plus step commutes with a constant shift of h, so the
// decisions are those of the normalised chain up to rounding, and u lies in [w min M, w max M]
// whatever the magnitudes of h.  The normalisation, the argmin and the stores then hang off the
// step-to-step dependency chain (h -> LDS -> candidates -> minimum -> u) instead of lying on
// it.  The shifted frame differs from the normalised one by a per-scanline constant k (k = 0 at
// the first sender, k' = lse - shift per step), so the stored message msg - lse needs no
// correction and scale = exp(lse - k).
// Planes are addressed through buffer resources (every [H,W,C] plane below 2^31 bytes) with one
// scalar offset per position; lanes past C or past the last scanline store to an offset past
// the resource (dropped), so the steps have no branches.
// ---------------------------------------------------------------------------------------
template <int CM, bool VERT, int PF>
__global__ __launch_bounds__(64) void lbp_fwd_semg(FwdArgs a) {
  constexpr int SEG = 32, SPW = 2;
  __shared__ float lds_h[SPW][SEG];  // a row per segment: every lane writes its slot
  const int H = a.H, W = a.W, C = a.C;
  const int L = VERT ? W : H;
  const int len = VERT ? H : W;
  const int nsteps = len - 1;
  const int wpg = (L + SPW - 1) / SPW;
  // columns of a vertical launch go to the XCDs in contiguous ranges (shared cache lines)
  const int bi = VERT ? xcd_contiguous_block((int)blockIdx.x, (int)gridDim.x) : (int)blockIdx.x;
  const int g = bi / wpg;
  const int wl = bi - g * wpg;
  const int n = g >> 1;
  const int dir = VERT ? ((g & 1) ? UP : DOWN) : ((g & 1) ? LEFT : RIGHT);
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
  const int cl = lab ? c : C - 1;  // label loaded by this lane
  const int64_t HW = (int64_t)H * W;
  const int64_t HWC = HW * C;
  // pixel at position p of the scanline: lbase + p * pstride.  Wave-uniform byte offsets of a
  // position in planes of C floats, C bytes and one float per pixel; the scalar offsets of the
  // sender and receiver positions advance by a signed step per sender (always in range).
  const int lbase = VERT ? ln : ln * W;
  const uint32_t pstride = VERT ? (uint32_t)W : 1u;
  const uint32_t st_f = pstride * (uint32_t)C * 4u, st_b = pstride * (uint32_t)C, st_p = pstride * 4u;
  const uint32_t p_first = fwd ? 0u : (uint32_t)nsteps;
  const uint32_t d_f = fwd ? st_f : 0u - st_f, d_b = fwd ? st_b : 0u - st_b, d_p = fwd ? st_p : 0u - st_p;
  auto pos = [&](int s) -> uint32_t { return (uint32_t)(fwd ? s : nsteps - s); };
  // per-lane byte offsets of the scanline's pixel 0: loads, stores (dropped past C or past L)
  const uint32_t v_ld = (uint32_t)(lbase * C + cl) * 4u;
  const uint32_t v_st = st_ok ? (uint32_t)(lbase * C + c) * 4u : kBufDrop;
  const uint32_t v_stb = st_ok ? (uint32_t)(lbase * C + c) : kBufDrop;
  const uint32_t v_e = (uint32_t)lbase * 4u;
  const uint32_t v_pp = (line_ok && c == 0) ? (uint32_t)lbase * 4u : kBufDrop;
  const uint32_t plane = (uint32_t)(HWC * 4);
  const float *xb = VERT ? a.msg_out : a.msg_in;
  const auto cost_r = make_rsrc(a.cost + (int64_t)n * HWC, plane);
  const auto x1_r = make_rsrc(xb + ((int64_t)n * 4 + (VERT ? RIGHT : UP)) * HWC, plane);
  const auto x2_r = make_rsrc(xb + ((int64_t)n * 4 + (VERT ? LEFT : DOWN)) * HWC, plane);
  const auto mi_r = make_rsrc(a.msg_in + ((int64_t)n * 4 + rslot) * HWC, plane);
  const auto out_r = make_rsrc(a.msg_out + ((int64_t)n * 4 + rslot) * HWC, plane);
  const auto arg_r = a.argmin ? make_rsrc(a.argmin + ((int64_t)n * 4 + rslot) * HWC, plane / 4u)
                              : make_rsrc(a.msg_out, 0u);
  const auto edge_r = make_rsrc(a.edge + ((int64_t)n * 4 + dir) * HW, (uint32_t)HW * 4u);
  const auto sc_r = a.scale ? make_rsrc(a.scale + ((int64_t)n * 4 + rslot) * HW, (uint32_t)HW * 4u)
                            : make_rsrc(a.msg_out, 0u);

  // this label's column of the direction's matrix
  float Mc[CM];
#pragma unroll
  for (int l = 0; l < CM; ++l) Mc[l] = (l < C && lab) ? a.jump[((int64_t)dir * C + l) * C + c] : 0.f;

  // the label of the minimum is 0 in this branch at every position of the scanline
  if (a.gargmin) {
    const auto ga_r = make_rsrc(a.gargmin + ((int64_t)n * 4 + rslot) * HW, (uint32_t)HW);
    for (int p = c; p < len; p += SEG)
      buf_st8(ga_r, line_ok ? (uint32_t)lbase + (uint32_t)p * pstride : kBufDrop, 0u, 0);
  }
  // the first sender's chain slot is never written by this pass: it keeps the input message
  float u = buf_ld(mi_r, v_ld, p_first * st_f);
  buf_st(out_r, v_st, p_first * st_f, u);
  buf_st8(arg_r, v_stb, p_first * st_b, 0);
  buf_st(sc_r, v_pp, p_first * st_p, 0.f);
  float k = 0.f;  // shifted frame minus the normalised frame, per scanline

  float pc[PF], p1[PF], p2[PF], pw[PF];
  auto load_at = [&](uint32_t of, uint32_t op, int i) {
    pc[i] = buf_ld(cost_r, v_ld, of);
    p1[i] = buf_ld(x1_r, v_ld, of);
    p2[i] = buf_ld(x2_r, v_ld, of);
    pw[i] = buf_ld(edge_r, v_e, op);
  };
  auto load_step = [&](int s, int i) { load_at(pos(s) * st_f, pos(s) * st_p, i); };
#pragma unroll
  for (int i = 0; i < PF; ++i) {
    load_step(i < nsteps ? i : (nsteps > 0 ? nsteps - 1 : 0), i);
    __builtin_amdgcn_sched_barrier(0);  // keep the ring slots in issue order
  }
  // receiver offsets of the next step
  uint32_t o_f = p_first * st_f + d_f, o_b = p_first * st_b + d_b, o_p = p_first * st_p + d_p;
  auto step = [&](int i) {
    // the chain: h -> LDS -> candidates -> minimum -> u
    const float h = ((pc[i] + p1[i]) + p2[i]) + u;
    const float hl = lab ? h : __builtin_inff();
    lds_h[seg][c] = hl;
    wave_lds_sync();
    // the shift min_l h[l], over the lanes while the row makes its LDS round trip
    const float shift = fminf(seg_reduce<SEG>(hl, OpMin()), 1e15f);  // all-infinite h: the clamp
    const float w = pw[i];
    float hv[CM], nv[CM];
#pragma unroll
    for (int l = 0; l < CM; ++l) hv[l] = lds_h[seg][l];
    wave_lds_sync();  // the reads are done before the next step's write
#pragma unroll
    for (int l = 0; l < CM; ++l) nv[l] = fmaf(Mc[l], w, hv[l]);
    // minimum: a min3 tree over the candidates and the initial 1e15
    float m[CM + 1];
#pragma unroll
    for (int l = 0; l < CM; ++l) m[l] = nv[l];
    m[CM] = 1e15f;
    int cnt = CM + 1;
#pragma unroll
    for (int r = 0; r < 6; ++r) {
      if (cnt > 1) {
        int o = 0;
#pragma unroll
        for (int t = 0; t < CM + 1; t += 3) {
          if (t < cnt) {
            float v = m[t];
            if (t + 1 < cnt) v = fminf(v, m[t + 1]);
            if (t + 2 < cnt) v = fminf(v, m[t + 2]);
            m[o++] = v;
          }
        }
        cnt = o;
      }
    }
    const float msg = m[0];
    u = msg - shift;
    // off the chain: normalisation, argmin (the last label equal to the minimum), stores
    const float mx = seg_reduce<SEG>(lab ? msg : -__builtin_inff(), OpMax());
    const float se = seg_reduce<SEG>(lab ? fast_exp(msg - mx) : 0.f, OpSum());
    const float lse = mx + fast_log(se);
    // (a chain of selects in label order: the scheduler interleaves it with the next step's
    // chain; a max tree over per-label candidates took more instructions)
    int arg = 0;
#pragma unroll
    for (int l = 0; l < CM; ++l) arg = nv[l] == msg ? l : arg;
    buf_st(out_r, v_st, o_f, msg - lse);
    buf_st8(arg_r, v_stb, o_b, (uint8_t)arg);
    buf_st(sc_r, v_pp, o_p, fast_exp(lse - k));
    k = lse - shift;
    o_f += d_f;
    o_b += d_b;
    o_p += d_p;
  };
  // main loop: the ring refills stay within the scanline (s + PF <= nsteps - 1)
  int s0 = 0;
  uint32_t l_f = pos(PF < nsteps ? PF : 0) * st_f, l_p = pos(PF < nsteps ? PF : 0) * st_p;
  for (; s0 + 2 * PF <= nsteps; s0 += PF) {
#pragma unroll
    for (int i = 0; i < PF; ++i) {
      step(i);
      // refill ring slot i (its old contents are dead now: no register copies, no drain)
      load_at(l_f, l_p, i);
      l_f += d_f;
      l_p += d_p;
    }
  }
  // the last groups: refills clamped to the last sender, steps past the end skipped
  for (; s0 < nsteps; s0 += PF) {
#pragma unroll
    for (int i = 0; i < PF; ++i) {
      if (s0 + i < nsteps) {
        step(i);
        load_step(min(s0 + i + PF, nsteps - 1), i);
      }
    }
  }
}

template <int CM>
void launch_semg(const FwdArgs &a, bool vert, hipStream_t st) {
  const int L = vert ? a.W : a.H;
  const dim3 grid((unsigned)(a.N * 2 * ((L + 1) / 2)));
  // an 8-step ring (4 steps: 0.413 / 0.263 ms instead of 0.400 / 0.253 at 2x512x1024x19)
  if (vert) hipLaunchKernelGGL((lbp_fwd_semg<CM, true, 8>), grid, dim3(kWave), 0, st, a);
  else hipLaunchKernelGGL((lbp_fwd_semg<CM, false, 8>), grid, dim3(kWave), 0, st, a);
}


// ---------------------------------------------------------------------------------------
// Semantic per-pixel pairwise, full min-plus branch (delta >= C), C <= CM (20 or 32): the
// lane layout and step of lbp_fwd_semg (shifted frame, branch-free buffer-addressed steps), with
// the sender's matrix column M_q[dir][.][c] read per step into the prefetch ring: CM loads that
// share the step's scalar offset, each with a per-lane offset computed once (row l of the
// sender's C x C block), so a step carries no address arithmetic for them.  IMGS: the two
// half-waves are the two images of the same scanline (N == 2), so both read the same matrix
// addresses in one instruction; otherwise two scanlines of one image.  Tensors are addressed
// as whole buffers (the messages of all N images below 2^31 bytes, a direction's matrix plane
// too), the image and slot parts in the per-lane offsets.
// ---------------------------------------------------------------------------------------
template <int CM, bool VERT, int PF, bool IMGS>
__global__ __launch_bounds__(64) void lbp_fwd_semp(FwdArgs a) {
  constexpr int SEG = 32, SPW = 2;
  __shared__ float lds_h[SPW][SEG];  // a row per segment: every lane writes its slot
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
  const int g = bi / wpg;
  const int wl = bi - g * wpg;
  const int n = IMGS ? seg : g >> 1;
  const int dir = VERT ? ((g & 1) ? UP : DOWN) : ((g & 1) ? LEFT : RIGHT);
  const int rslot = dir == RIGHT ? LEFT : dir == LEFT ? RIGHT : dir == DOWN ? UP : DOWN;
  const bool fwd = (dir == RIGHT || dir == DOWN);
  const int line = IMGS ? wl : wl * SPW + seg;
  const bool line_ok = line < L;
  const int ln = line_ok ? line : 0;
  const bool lab = c < C;
  const bool st_ok = line_ok && lab;
  const int cl = lab ? c : C - 1;  // label loaded by this lane
  const uint32_t HWu = (uint32_t)H * (uint32_t)W, HWCu = HWu * (uint32_t)C;
  // pixel at position p of the scanline: lbase + p * pstride; wave-uniform byte offsets of a
  // position in planes of C floats, C bytes, one float per pixel and C x C floats
  const uint32_t lbase = VERT ? (uint32_t)ln : (uint32_t)ln * (uint32_t)W;
  const uint32_t pstride = VERT ? (uint32_t)W : 1u;
  const uint32_t st_f = pstride * (uint32_t)C * 4u, st_b = pstride * (uint32_t)C, st_p = pstride * 4u;
  const uint32_t st_m = st_f * (uint32_t)C;
  const uint32_t p_first = fwd ? 0u : (uint32_t)nsteps;
  const uint32_t d_f = fwd ? st_f : 0u - st_f, d_b = fwd ? st_b : 0u - st_b, d_p = fwd ? st_p : 0u - st_p;
  const uint32_t d_m = fwd ? st_m : 0u - st_m;
  auto pos = [&](int s) -> uint32_t { return (uint32_t)(fwd ? s : nsteps - s); };
  // per-lane byte offsets of the scanline's pixel 0 in the image's planes
  const uint32_t lpix = lbase * (uint32_t)C;
  auto slot_f = [&](int sl, int lbl) { return (((uint32_t)n * 4u + (uint32_t)sl) * HWCu + lpix + (uint32_t)lbl) * 4u; };
  const uint32_t v_cost = ((uint32_t)n * HWCu + lpix + (uint32_t)cl) * 4u;
  const uint32_t v_x1 = slot_f(VERT ? RIGHT : UP, cl), v_x2 = slot_f(VERT ? LEFT : DOWN, cl);
  const uint32_t v_mi = slot_f(rslot, cl);
  const uint32_t v_st = st_ok ? slot_f(rslot, c) : kBufDrop;
  const uint32_t v_stb = st_ok ? ((uint32_t)n * 4u + (uint32_t)rslot) * HWCu + lpix + (uint32_t)c : kBufDrop;
  const uint32_t v_e = (((uint32_t)n * 4u + (uint32_t)dir) * HWu + lbase) * 4u;
  const uint32_t v_pp = (line_ok && c == 0) ? (((uint32_t)n * 4u + (uint32_t)rslot) * HWu + lbase) * 4u : kBufDrop;
  const uint32_t N4 = (uint32_t)a.N * 4u;
  const float *xb = VERT ? a.msg_out : a.msg_in;
  const auto cost_r = make_rsrc(a.cost, (uint32_t)a.N * HWCu * 4u);
  const auto x_r = make_rsrc(xb, N4 * HWCu * 4u);
  const auto mi_r = make_rsrc(a.msg_in, N4 * HWCu * 4u);
  const auto out_r = make_rsrc(a.msg_out, N4 * HWCu * 4u);
  const auto arg_r = a.argmin ? make_rsrc(a.argmin, N4 * HWCu) : make_rsrc(a.msg_out, 0u);
  const auto edge_r = make_rsrc(a.edge, N4 * HWu * 4u);
  const auto sc_r = a.scale ? make_rsrc(a.scale, N4 * HWu * 4u) : make_rsrc(a.msg_out, 0u);
  const auto mat_r = make_rsrc(a.jump + (int64_t)dir * HWu * C * C, HWCu * (uint32_t)C * 4u);
  // row l of the sender's C x C block, label cl (rows past C read a neighbouring block or
  // past the plane, whose values meet h = +inf only)
  uint32_t v_m[CM];
#pragma unroll
  for (int l = 0; l < CM; ++l) v_m[l] = ((lpix + (uint32_t)l) * (uint32_t)C + (uint32_t)cl) * 4u;

  // the label of the minimum is 0 in this branch at every position of the scanline
  if (a.gargmin) {
    const auto ga_r = make_rsrc(a.gargmin, N4 * HWu);
    const uint32_t gb = ((uint32_t)n * 4u + (uint32_t)rslot) * HWu + lbase;
    for (int p = c; p < len; p += SEG) buf_st8(ga_r, line_ok ? gb + (uint32_t)p * pstride : kBufDrop, 0u, 0);
  }
  // the first sender's chain slot is never written by this pass: it keeps the input message
  float u = buf_ld(mi_r, v_mi, p_first * st_f);
  buf_st(out_r, v_st, p_first * st_f, u);
  buf_st8(arg_r, v_stb, p_first * st_b, 0);
  buf_st(sc_r, v_pp, p_first * st_p, 0.f);
  float k = 0.f;  // shifted frame minus the normalised frame, per scanline

  float pc[PF], p1[PF], p2[PF], pw[PF], pm[PF][CM];
  auto load_at = [&](uint32_t of, uint32_t op, uint32_t om, int i) {
    pc[i] = buf_ld(cost_r, v_cost, of);
    p1[i] = buf_ld(x_r, v_x1, of);
    p2[i] = buf_ld(x_r, v_x2, of);
    pw[i] = buf_ld(edge_r, v_e, op);
#pragma unroll
    for (int l = 0; l < CM; ++l) pm[i][l] = buf_ld<BPL_SEMP_AUX>(mat_r, v_m[l], om);
  };
  auto load_step = [&](int s, int i) { load_at(pos(s) * st_f, pos(s) * st_p, pos(s) * st_m, i); };
#pragma unroll
  for (int i = 0; i < PF; ++i) {
    load_step(i < nsteps ? i : (nsteps > 0 ? nsteps - 1 : 0), i);
    __builtin_amdgcn_sched_barrier(0);  // keep the ring slots in issue order
  }
  // receiver offsets of the next step
  uint32_t o_f = p_first * st_f + d_f, o_b = p_first * st_b + d_b, o_p = p_first * st_p + d_p;
  auto step = [&](int i) {
    // the chain: h -> LDS -> candidates -> minimum -> u
    const float h = ((pc[i] + p1[i]) + p2[i]) + u;
    const float hl = lab ? h : __builtin_inff();
    lds_h[seg][c] = hl;
    wave_lds_sync();
    // the shift min_l h[l], over the lanes while the row makes its LDS round trip
    const float shift = fminf(seg_reduce<SEG>(hl, OpMin()), 1e15f);  // all-infinite h: the clamp
    const float w = pw[i];
    float hv[CM], nv[CM];
#pragma unroll
    for (int l = 0; l < CM; ++l) hv[l] = lds_h[seg][l];
    wave_lds_sync();  // the reads are done before the next step's write
#pragma unroll
    for (int l = 0; l < CM; ++l) nv[l] = fmaf(pm[i][l], w, hv[l]);
    // minimum: a min3 tree over the candidates and the initial 1e15
    float m[CM + 1];
#pragma unroll
    for (int l = 0; l < CM; ++l) m[l] = nv[l];
    m[CM] = 1e15f;
    int cnt = CM + 1;
#pragma unroll
    for (int r = 0; r < 6; ++r) {
      if (cnt > 1) {
        int o = 0;
#pragma unroll
        for (int t = 0; t < CM + 1; t += 3) {
          if (t < cnt) {
            float v = m[t];
            if (t + 1 < cnt) v = fminf(v, m[t + 1]);
            if (t + 2 < cnt) v = fminf(v, m[t + 2]);
            m[o++] = v;
          }
        }
        cnt = o;
      }
    }
    const float msg = m[0];
    u = msg - shift;
    // off the chain: normalisation, argmin (the last label equal to the minimum), stores
    const float mx = seg_reduce<SEG>(lab ? msg : -__builtin_inff(), OpMax());
    const float se = seg_reduce<SEG>(lab ? fast_exp(msg - mx) : 0.f, OpSum());
    const float lse = mx + fast_log(se);
    // (a chain of selects in label order: the scheduler interleaves it with the next step's
    // chain; a max tree over per-label candidates took more instructions)
    int arg = 0;
#pragma unroll
    for (int l = 0; l < CM; ++l) arg = nv[l] == msg ? l : arg;
    buf_st(out_r, v_st, o_f, msg - lse);
    buf_st8(arg_r, v_stb, o_b, (uint8_t)arg);
    buf_st(sc_r, v_pp, o_p, fast_exp(lse - k));
    k = lse - shift;
    o_f += d_f;
    o_b += d_b;
    o_p += d_p;
  };
  // main loop: the ring refills stay within the scanline (s + PF <= nsteps - 1)
  int s0 = 0;
  {
    const uint32_t pr = pos(PF < nsteps ? PF : 0);
    uint32_t l_f = pr * st_f, l_p = pr * st_p, l_m = pr * st_m;
    for (; s0 + 2 * PF <= nsteps; s0 += PF) {
#pragma unroll
      for (int i = 0; i < PF; ++i) {
        step(i);
        // refill ring slot i (its old contents are dead now: no register copies, no drain)
        load_at(l_f, l_p, l_m, i);
        l_f += d_f;
        l_p += d_p;
        l_m += d_m;
      }
    }
  }
  // the last groups: refills clamped to the last sender, steps past the end skipped
  for (; s0 < nsteps; s0 += PF) {
#pragma unroll
    for (int i = 0; i < PF; ++i) {
      if (s0 + i < nsteps) {
        step(i);
        load_step(min(s0 + i + PF, nsteps - 1), i);
      }
    }
  }
}

template <int CM, int PF>
void launch_semp(const FwdArgs &a, bool vert, hipStream_t st) {
  const int L = vert ? a.W : a.H;
  const bool imgs = a.N == 2;
  const dim3 grid((unsigned)(imgs ? 2 * L : a.N * 2 * ((L + 1) / 2)));
  if (imgs) {
    if (vert) hipLaunchKernelGGL((lbp_fwd_semp<CM, true, PF, true>), grid, dim3(kWave), 0, st, a);
    else hipLaunchKernelGGL((lbp_fwd_semp<CM, false, PF, true>), grid, dim3(kWave), 0, st, a);
  } else {
    if (vert) hipLaunchKernelGGL((lbp_fwd_semp<CM, true, PF, false>), grid, dim3(kWave), 0, st, a);
    else hipLaunchKernelGGL((lbp_fwd_semp<CM, false, PF, false>), grid, dim3(kWave), 0, st, a);
  }
}

// ---------------------------------------------------------------------------------------
// Stereo fast path for 16 < C <= 256: one scanline per wave, label c = lane * LPL + k, the
// candidate window from neighbouring lanes by DPP shifts (no LDS round trip), and no
// per-label control flow in the step:
//   * labels past C (only when C != 64 * LPL) carry h = +inf, so they never win a minimum,
//     and the neighbour-lane shifts fill the wave's ends with +inf, so window candidates
//     outside [0, C) need no range test;
//   * minima are reduced as order-preserving integer keys and sums as floats, with the DPP
//     row stages folded into the arithmetic and the half/row exchanges done by permlane
//     swaps;
//   * rows are addressed as a uniform scanline pointer plus a per-lane 32-bit offset, and a
//     full label row is one vector load/store per lane.
// The message, argmin and normalisation semantics are those of lbp_fwd_phase.
// ---------------------------------------------------------------------------------------
constexpr int kArgTrunc = 64;  // window-offset sentinel: "truncation term won"

template <int LPL>
__device__ __forceinline__ void ld_lanes(const float *row, uint32_t voff, float *d) {
  const float *p = reinterpret_cast<const float *>(reinterpret_cast<const char *>(row) + voff);
#pragma unroll
  for (int k = 0; k < LPL; ++k) d[k] = p[k];
}
template <int LPL>
__device__ __forceinline__ void st_lanes(float *row, uint32_t voff, const float *d) {
  float *p = reinterpret_cast<float *>(reinterpret_cast<char *>(row) + voff);
#pragma unroll
  for (int k = 0; k < LPL; ++k) p[k] = d[k];
}
__device__ __forceinline__ float shr1_inf(float v) {  // lane i <- lane i-1, lane 0 <- +inf
  return __int_as_float(__builtin_amdgcn_update_dpp(0x7F800000, __float_as_int(v), 0x138, 0xF, 0xF, false));
}
__device__ __forceinline__ float shl1_inf(float v) {  // lane i <- lane i+1, lane 63 <- +inf
  return __int_as_float(__builtin_amdgcn_update_dpp(0x7F800000, __float_as_int(v), 0x130, 0xF, 0xF, false));
}

// Fast step (FASTK = K in {6, 8}; launched for delta == 3 with argmin, gargmin and scale
// all requested (OUTS) or none of them, and used for every scanline whose window offsets
// are all zero).  The step has
// no branches except rare uniform ones placed after the next pixel's chain value, so the
// compiler can overlap the tail of one step (label of the minimum, normalisation, stores)
// with the dependent chain of the next:
//   * the chain carried to the next pixel is u = msg - gv instead of the normalised message
//     msg - lse.  Messages are shift-invariant in their input (a constant added to h moves
//     every candidate and gv by that constant), so this changes no message and no argmin; the
//     normalised output is u - L with L = log sum exp(u), and the scale exp(lse) of the true
//     frame is exp(gv + L - L_prev), L_prev being the previous pixel's L (0 at the start);
//   * the window candidates are formed before the global minimum is known; their minimum is a
//     min3 tree and the winner the last offset attaining it (the reference's `<=` scan); gv
//     precedes the window in the reference's candidate order, so a window value equal to gv
//     wins the merge;
//   * the label of the minimum comes from ballots (scalar ffs); ties, which are rare, take the
//     largest tie key (an integer max-reduction), the reference's tree order;
//   * the log-sum-exp needs no shift: every u lies in [(Fmin - JK) w, 0], so unless
//     (JK - Fmin) w > 80 no exp(u) underflows; otherwise (rare) it is shifted by max u.
// Sweep of the P1/P2 form: jump length 6 with J[1] == J[5] == JK (the truncation value) and
// J[3] == 0 at every pixel of the scanline, zero window offsets, delta 3 -- the packing of
// per-pixel (P1-, P1+, P2) affinities (message_passing_op_cuda.py:35-56) -- and no
// log-sum-exp term that can underflow ((JK - J[j]) w <= 80), all checked by the caller.
//   * The window candidates at offsets -3, -2, 2, 3 carry the truncation value JK, so they are
//     never below the global term gv = JK w + min h, and equal it exactly when their sender
//     label attains min h: they only decide the argmin at such ties.  The window minimum is
//     therefore min(J[4] w + h[c-1], h[c], J[2] w + h[c+1]) and the message min(that, gv).
//   * The argmin is the last candidate attaining the message in the reference's order (gv,
//     then offsets -3 .. 3).  An offset +-2 / +-3 attains it exactly when the message is gv
//     and its sender attains min h; its sender label is then the global argmin itself.  (This
//     is the FP64 semantics of those ties; an fp32 fused multiply-add of a slightly larger h
//     could also round onto gv.)  With several labels attaining min h (rare) every one is
//     tested.
//   * The chain carried to the next pixel is u = msg - gv (messages are shift-invariant in h);
//     every u lies in [-80, 0], so the log-sum-exp needs no shift.
//   * Row streams are buffer-addressed: one wave-uniform row offset shared by the cost and
//     both cross-message streams (their planes are per-lane offsets of one resource per
//     tensor); ring loads past the last sender run off the scanline and are never used.
//   * BV (vertical): the cost and both horizontal messages come as one base plane,
//     (cost + M_R) + M_L, so h = base + chain is the vertical order's sum bit for bit.
//   * PAIR (horizontal, both passes of the row in one workgroup): steps from Tb on (after a
//     workgroup barrier; Tb >= (W - 2) / 2, a multiple of PF) are second visits, whose receiver
//     the other pass has already reached: they write the receiver's base plane entry from
//     their own message, the other pass's stored message and the receiver's cost (the next
//     ring slot).  The caller writes the entries of the row ends and of the pixels both passes
//     reach before the barrier.
// cache policy of the P1/P2 sweeps' message and argmin stores, per phase: non-temporal (2),
// since nothing reads them before the next phase has streamed its own planes (same-run A/B at
// 384x1248x192: forward 0.806 / 0.480 -> 0.796 / 0.474 ms, vertical backward 0.590 -> 0.584 ms)
#ifndef BPL_FWD_ST_AUX_H
#define BPL_FWD_ST_AUX_H 2
#endif
#ifndef BPL_FWD_ST_AUX_V
#define BPL_FWD_ST_AUX_V 2
#endif
// cache policy of their row loads, per phase: default.  Non-temporal loads speed the horizontal
// phase up (0.794 -> 0.775 ms) but slow the vertical one that follows it, which reads the base
// plane the horizontal phase has just written, by more (0.474 -> 0.503 ms; same-run A/B at
// 384x1248x192), and slow it directly when used there (0.474 -> 0.493 ms); for the cost plane
// alone: horizontal 0.795 -> 0.789 ms, vertical 0.475 -> 0.478 ms, 0.1% per step (not kept)
#ifndef BPL_FWD_LD_AUX_H
#define BPL_FWD_LD_AUX_H 0
#endif
#ifndef BPL_FWD_LD_AUX_V
#define BPL_FWD_LD_AUX_V 0
#endif
template <int LPL, bool VERT, bool FULL, bool OUTS, int PF, bool BV = false, bool PAIR = false>
__device__ __forceinline__ void fwd_sweep_p1p2(const FwdArgs &a, int n, int dir, int rslot, int nj, int64_t q0,
                                               int64_t dq, int nsteps, int lane, int c0, const bool *valid,
                                               int Tb = 0) {
  static_assert(!(BV && !VERT) && !(PAIR && VERT), "base plane: read by vertical, written by horizontal sweeps");
  constexpr int kMsgAux = VERT ? BPL_FWD_ST_AUX_V : BPL_FWD_ST_AUX_H;  // messages' and argmins' store policy
  constexpr int kLdAux = VERT ? BPL_FWD_LD_AUX_V : BPL_FWD_LD_AUX_H;  // row loads' policy
  constexpr int ND = (3 + LPL - 1) / LPL;
  const int H = a.H, W = a.W, C = a.C, K = a.K;
  const int64_t HW = (int64_t)H * W;
  const uint32_t pb = (uint32_t)(HW * C * 4);        // bytes of one [H,W,C] plane
  const int64_t img4 = (int64_t)n * 4 * HW * C;      // the image's 4-slot message block
  const __amdgpu_buffer_rsrc_t c_rs = make_rsrc(a.cost + (int64_t)n * HW * C, pb);
  const __amdgpu_buffer_rsrc_t mi_rs = make_rsrc(a.msg_in + img4, a.in_zero ? 0u : 4u * pb);
  const __amdgpu_buffer_rsrc_t mo_rs = make_rsrc(a.msg_out + img4, 4u * pb);
  const __amdgpu_buffer_rsrc_t x_rs = VERT ? mo_rs : mi_rs;     // the cross messages
  const int64_t rs_base = ((int64_t)n * 4 + rslot) * HW;
  const __amdgpu_buffer_rsrc_t j_rs = make_rsrc(a.jump + ((int64_t)nj * 4 + dir) * K * HW, (uint32_t)(K * HW * 4));
  const __amdgpu_buffer_rsrc_t e_rs = make_rsrc(a.edge + ((int64_t)n * 4 + dir) * HW, (uint32_t)(HW * 4));
  const void *a_base = OUTS ? (const void *)(a.argmin + rs_base * C) : (const void *)a.cost;
  const void *s_base = OUTS ? (const void *)(a.scale + rs_base) : (const void *)a.cost;
  const void *g_base = OUTS ? (const void *)(a.gargmin + rs_base) : (const void *)a.cost;
  const __amdgpu_buffer_rsrc_t a_rs = make_rsrc(a_base, OUTS ? pb / 4u : 0u);
  const __amdgpu_buffer_rsrc_t b_rs = make_rsrc((BV || PAIR) ? (const void *)(a.base + (int64_t)n * HW * C)
                                                             : (const void *)a.cost, (BV || PAIR) ? pb : 0u);
  const __amdgpu_buffer_rsrc_t s_rs = make_rsrc(s_base, OUTS ? (uint32_t)(HW * 4) : 0u);
  const __amdgpu_buffer_rsrc_t g_rs = make_rsrc(g_base, OUTS ? (uint32_t)HW : 0u);
  // per-lane byte offsets: the lane's labels within a row, plus the plane of the stream
  uint32_t vl[LPL];
#pragma unroll
  for (int k = 0; k < LPL; ++k) vl[k] = (uint32_t)(FULL ? c0 + k : min(c0 + k, C - 1)) * 4u;
  const uint32_t p_x1 = (uint32_t)(VERT ? RIGHT : UP) * pb, p_x2 = (uint32_t)(VERT ? LEFT : DOWN) * pb;
  const uint32_t p_o = (uint32_t)rslot * pb;
  const uint32_t jv = lane < K ? (uint32_t)lane * (uint32_t)(HW * 4) : kBufDrop;  // lane j: jump entry j
  const uint32_t ev = lane == K ? 0u : kBufDrop;                                   // lane K: edge weight
  const uint32_t one = lane == 0 ? 0u : kBufDrop;   // per-receiver values: one lane stores
  const uint32_t rowb = (uint32_t)C * 4u;
  const uint32_t drow = (uint32_t)((int)dq * (int)rowb);    // row offset step (two's complement)
  const uint32_t dpix = (uint32_t)((int)dq * 4);
  auto ld_row = [&](__amdgpu_buffer_rsrc_t r, uint32_t plane, uint32_t soff, float *d) {
    if (FULL) {
      buf_ld_row<LPL, kLdAux>(r, vl[0] + plane, soff, d);
    } else {
#pragma unroll
      for (int k = 0; k < LPL; ++k) d[k] = buf_ld<kLdAux>(r, vl[k] + plane, soff);
    }
  };
  auto st_plane = [&](__amdgpu_buffer_rsrc_t r, uint32_t plane, uint32_t soff, const float *v) {
    if (FULL) {
      buf_st_row<LPL>(r, vl[0] + plane, soff, v);
    } else {
#pragma unroll
      for (int k = 0; k < LPL; ++k) buf_st(r, valid[k] ? vl[k] + plane : kBufDrop, soff, v[k]);
    }
  };
  // the messages written by the sweep (cache policy kMsgAux)
  auto st_row = [&](uint32_t soff, const float *v) {
    if (FULL) {
      buf_st_row<LPL, kMsgAux>(mo_rs, vl[0] + p_o, soff, v);
    } else {
#pragma unroll
      for (int k = 0; k < LPL; ++k) buf_st<kMsgAux>(mo_rs, valid[k] ? vl[k] + p_o : kBufDrop, soff, v[k]);
    }
  };
  auto st_args = [&](uint32_t soff_b, const uint32_t *b) {   // soff_b: byte offset of the row
    if constexpr (FULL && LPL == 3) {
      __builtin_amdgcn_raw_buffer_store_b16((uint16_t)(b[0] | (b[1] << 8)), a_rs, (uint32_t)c0, soff_b, kMsgAux);
      __builtin_amdgcn_raw_buffer_store_b8((uint8_t)b[2], a_rs, (uint32_t)c0 + 2u, soff_b, kMsgAux);
    } else if constexpr (FULL && LPL == 2) {
      __builtin_amdgcn_raw_buffer_store_b16((uint16_t)(b[0] | (b[1] << 8)), a_rs, (uint32_t)c0, soff_b, kMsgAux);
    } else {
#pragma unroll
      for (int k = 0; k < LPL; ++k)
        __builtin_amdgcn_raw_buffer_store_b8((uint8_t)b[k], a_rs, valid[k] ? (uint32_t)(c0 + k) : kBufDrop, soff_b,
                                             kMsgAux);
    }
  };

  // the first sender's chain is its input message, copied through
  const uint32_t so0 = (uint32_t)q0 * rowb;
  float chain[LPL];
  ld_row(mi_rs, p_o, so0, chain);
  st_row(so0, chain);
  if constexpr (OUTS) {
    uint32_t z[LPL];
#pragma unroll
    for (int k = 0; k < LPL; ++k) z[k] = 0u;
    st_args(so0 / 4u, z);
    buf_st(s_rs, one, (uint32_t)q0 * 4u, 0.f);
    __builtin_amdgcn_raw_buffer_store_b8((uint8_t)0, g_rs, one, (uint32_t)q0, 0);
  }
  // labels of this lane and their neighbours (the argmin candidates c-1, c, c+1)
  uint32_t lab_c[LPL + 2];
#pragma unroll
  for (int k = 0; k < LPL + 2; ++k) lab_c[k] = (uint32_t)(c0 + k - 1);

  float pc[PF][LPL], q1[PF][LPL], q2[PF][LPL], pj[PF], pe[PF];
  uint32_t ring_so = so0, ring_sp = (uint32_t)q0 * 4u;   // offsets of the next step to load
  auto load_ring = [&](float *dc, float *d1, float *d2, float &dj, float &de) {
    if constexpr (BV) {
      ld_row(b_rs, 0u, ring_so, dc);
    } else {
      ld_row(c_rs, 0u, ring_so, dc);
      ld_row(x_rs, p_x1, ring_so, d1);
      ld_row(x_rs, p_x2, ring_so, d2);
    }
    dj = buf_ld(j_rs, jv, ring_sp);
    de = buf_ld(e_rs, ev, ring_sp);
    ring_so += drow;
    ring_sp += dpix;
  };
  // Stores of one step: the vector-memory counter counts them with the loads, and the wait
  // counts of the loop are the smaller of those of its two entries (the prologue and the back
  // edge).  A prologue of loads alone would make every wait in the loop stricter than the ring
  // needs (the loads of a slot then must have returned three steps after their issue instead of
  // about PF - 1), so the prologue issues as many dropped stores between its slots as a step
  // issues between two refills.
  // (A resource of 0 bytes drops every store; distinct offsets keep the stores from being
  // merged, as overwrites of one location or as one wider access.)
  constexpr int kArgSt = FULL ? (LPL == 3 ? 2 : LPL == 2 ? 1 : LPL) : LPL;
  constexpr int kStepSt = (OUTS ? kArgSt + 2 : 0) + (FULL ? 1 : LPL);
  const __amdgpu_buffer_rsrc_t none_rs = make_rsrc(a.msg_out, 0u);
#pragma unroll
  for (int i = 0; i < PF; ++i) {
    load_ring(pc[i], q1[i], q2[i], pj[i], pe[i]);
    if (i + 1 < PF) {
#pragma unroll
      for (int d = 0; d < kStepSt; ++d) buf_st(none_rs, (uint32_t)(i * kStepSt + d) * 256u, 0u, 0.f);
    }
    __builtin_amdgcn_sched_barrier(0);
  }
  float Lprev = 0.f;
  uint32_t rso = so0 + drow, rsp = (uint32_t)q0 * 4u + dpix;   // receiver of step 0
  // PAIR: the other pass's message at the receiver of each ring step (second visits)
  float po[PAIR ? PF : 1][LPL];
  const uint32_t p_oth = (uint32_t)(dir == RIGHT ? RIGHT : LEFT) * pb;
  uint32_t oth_so = 0;
  auto load_other = [&](float *d) {
    ld_row(mo_rs, p_oth, oth_so, d);
    oth_so += drow;
  };
  auto step = [&](int i, auto second_tag) {
    constexpr bool SECOND = PAIR && decltype(second_tag)::value;
    float h[LPL];
#pragma unroll
    for (int k = 0; k < LPL; ++k) {
      float c = pc[i][k];
      if (VERT) asm("" : "+v"(c) : "v"(chain[k]));  // the vertical order adds the chain last
      const float v = BV ? c + chain[k]
                    : VERT ? ((c + q1[i][k]) + q2[i][k]) + chain[k] : ((c + chain[k]) + q1[i][k]) + q2[i][k];
      h[k] = valid[k] ? v : __builtin_inff();
    }
    float pv = pj[i] + pe[i];
    asm("" : "+v"(pv) : "v"(chain[0]));  // no instruction: orders pv after the chain
    const float w = readlane_f(pv, 6);
    const float JK = readlane_f(pv, 5);
    const float Jm = readlane_f(pv, 4);   // offset -1: f = J[K/2 + 1]
    const float Jp = readlane_f(pv, 2);   // offset +1: f = J[K/2 - 1]
    float hmin = h[0];
#pragma unroll
    for (int k = 1; k < LPL; ++k) hmin = fminf(hmin, h[k]);
    const float Hs = wave_min_f(hmin);
    const float hl = shr1_inf(h[LPL - 1]);   // h[c0 - 1]
    const float hr = shl1_inf(h[0]);         // h[c0 + LPL]
    const float gv = fmaf(JK, w, Hs);
    float cm[LPL], cp[LPL], msg[LPL];
#pragma unroll
    for (int k = 0; k < LPL; ++k) {
      cm[k] = fmaf(Jm, w, k == 0 ? hl : h[k > 0 ? k - 1 : 0]);
      cp[k] = fmaf(Jp, w, k == LPL - 1 ? hr : h[k < LPL - 1 ? k + 1 : 0]);
      msg[k] = fminf(fminf(fminf(cm[k], h[k]), cp[k]), gv);
      chain[k] = msg[k] - gv;
    }
    // ---- tail: argmins, normalisation, stores
    if constexpr (OUTS) {
      uint64_t mk[LPL];
      int cnt = 0;
#pragma unroll
      for (int k = 0; k < LPL; ++k) {
        mk[k] = __ballot(valid[k] && h[k] == Hs);
        cnt += __popcll(mk[k]);
      }
      uint32_t lab[LPL];
      int lstar = 0;
      if (cnt == 1) {
#pragma unroll
        for (int k = 0; k < LPL; ++k)
          if (mk[k]) lstar = (__ffsll((long long)mk[k]) - 1) * LPL + k;
        // receivers 2-3 labels above the global argmin tie through offsets -3 / -2 (before
        // the c-1, c, c+1 candidates), those 2-3 below through +2 / +3 (after them)
        const uint32_t above = (uint32_t)(lstar + 2), below = (uint32_t)(lstar - 3);
#pragma unroll
        for (int k = 0; k < LPL; ++k) {
          const bool atg = !(msg[k] < gv);   // the message is gv
          const uint32_t c = lab_c[k + 1];
          uint32_t l = atg && c - above <= 1u ? (uint32_t)lstar : (uint32_t)kFromTrunc;
          uint32_t hit = atg && c - below <= 1u ? 1u : 0u;
          // materialised before the selects (no instruction): otherwise the compiler guards
          // them with exec masks around the later overriding selects
          asm("" : "+v"(l), "+v"(hit));
          l = cm[k] == msg[k] ? lab_c[k] : l;
          l = h[k] == msg[k] ? c : l;
          l = cp[k] == msg[k] ? lab_c[k + 2] : l;
          lab[k] = hit ? (uint32_t)lstar : l;
        }
      } else {
        // several labels attain min h (rare): the tree order picks the stored one, and every
        // receiver whose offset-+-2/+-3 sender attains min h ties with gv
        int tk = -1;
#pragma unroll
        for (int k = 0; k < LPL; ++k) tk = max(tk, (valid[k] && h[k] == Hs) ? tie_key(c0 + k, a.bits) : -1);
        lstar = tie_key(wave_max_i(tk), a.bits);
        float Lh[ND][LPL], Rh[ND][LPL];
#pragma unroll
        for (int k = 0; k < LPL; ++k) { Lh[0][k] = shr1_inf(h[k]); Rh[0][k] = shl1_inf(h[k]); }
#pragma unroll
        for (int dd = 1; dd < ND; ++dd)
#pragma unroll
          for (int k = 0; k < LPL; ++k) { Lh[dd][k] = shr1_inf(Lh[dd - 1][k]); Rh[dd][k] = shl1_inf(Rh[dd - 1][k]); }
        auto hat = [&](int k, int kk) -> float {   // h at label c0 + k + kk, |kk| <= 3
          const int e = k + kk;
          if (e >= 0 && e < LPL) return h[e >= 0 && e < LPL ? e : 0];
          if (e < 0) {
            const int j = -e, dd = (j + LPL - 1) / LPL, slot = dd * LPL - j;
            return Lh[dd - 1][slot];
          }
          const int j = e - LPL + 1, dd = (j + LPL - 1) / LPL, slot = j - 1 - (dd - 1) * LPL;
          return Rh[dd - 1][slot];
        };
#pragma unroll
        for (int k = 0; k < LPL; ++k) {
          const bool atg = !(msg[k] < gv);
          const uint32_t c = lab_c[k + 1];
          uint32_t l = (uint32_t)kFromTrunc;
          l = atg && hat(k, -3) == Hs ? c - 3u : l;
          l = atg && hat(k, -2) == Hs ? c - 2u : l;
          l = cm[k] == msg[k] ? lab_c[k] : l;
          l = h[k] == msg[k] ? c : l;
          l = cp[k] == msg[k] ? lab_c[k + 2] : l;
          l = atg && hat(k, 2) == Hs ? c + 2u : l;
          lab[k] = atg && hat(k, 3) == Hs ? c + 3u : l;
        }
      }
      st_args(rso / 4u, lab);
      __builtin_amdgcn_raw_buffer_store_b8((uint8_t)lstar, g_rs, one, rsp / 4u, 0);
    }
    float se = 0.f;
#pragma unroll
    for (int k = 0; k < LPL; ++k) se += valid[k] ? fast_exp(chain[k]) : 0.f;
    se = wave_sum_f(se);
    const float Lc = fast_log(se);
    float outv[LPL];
#pragma unroll
    for (int k = 0; k < LPL; ++k) outv[k] = chain[k] - Lc;
    st_row(rso, outv);
    if constexpr (OUTS) buf_st(s_rs, one, rsp, fast_exp(gv + Lc - Lprev));
    if constexpr (SECOND) {
      // base entry of the receiver: its cost is the next ring slot's; the sums in the order
      // (cost + M_R) + M_L of the vertical h (this pass writes the L slot when it runs RIGHT)
      float bv[LPL];
#pragma unroll
      for (int k = 0; k < LPL; ++k) {
        float cn = pc[(i + 1) % PF][k], ot = po[i][k];
        asm("" : "+v"(cn), "+v"(ot) : "v"(outv[0]));
        bv[k] = dir == RIGHT ? (cn + ot) + outv[k] : (cn + outv[k]) + ot;
      }
      st_plane(b_rs, 0u, rso, bv);
    }
    Lprev = Lc;
    rso += drow;
    rsp += dpix;
  };
  auto group = [&](int i0, auto second_tag) {
#pragma unroll
    for (int i = 0; i < PF; ++i) {
      step(i, second_tag);
      load_ring(pc[i], q1[i], q2[i], pj[i], pe[i]);
      if constexpr (decltype(second_tag)::value) load_other(po[i]);
    }
  };
  int s0 = 0;
  if constexpr (PAIR) {
    if (Tb + PF <= nsteps - 1) {
      for (; s0 < Tb; s0 += PF) group(0, std::false_type());
      __syncthreads();  // the other pass's first visits are written
      oth_so = (uint32_t)(q0 + (int64_t)(Tb + 1) * dq) * rowb;
#pragma unroll
      for (int i = 0; i < PF; ++i) load_other(po[i]);
      for (; s0 + PF <= nsteps - 1; s0 += PF) group(0, std::true_type());
#pragma unroll
      for (int i = 0; i < PF; ++i) {
        if (s0 + i < nsteps - 1) step(i, std::true_type());
        else if (s0 + i == nsteps - 1) step(i, std::false_type());
      }
      return;
    }
  }
  for (; s0 + PF <= nsteps; s0 += PF) group(0, std::false_type());
#pragma unroll
  for (int i = 0; i < PF; ++i)
    if (s0 + i < nsteps) step(i, std::false_type());
}

// P1/P2 form of scanline `ln` of pass `dir` (wave-uniform): zero window offsets, J[1] == J[5]
// (the truncation value) and J[3] == 0 at every pixel, and no log-sum-exp term that can
// underflow, (JK - J[j]) w <= 80; the jump length is 6.
template <bool VERT>
__device__ __forceinline__ bool p1p2_scanline(const FwdArgs &a, int n, int dir, int ln, int lane) {
  const int H = a.H, W = a.W, K = a.K;
  const int len = VERT ? H : W;
  const int64_t HW = (int64_t)H * W;
  const int nj = a.Nj > 1 ? n : 0;
  const float *o_p = a.jump + ((int64_t)nj * 4 + dir) * K * HW;
  const float *e_p = a.edge + ((int64_t)n * 4 + dir) * HW;
  bool form = true, big = false;
  for (int p = lane; p < len; p += kWave) {
    const int64_t q = VERT ? (int64_t)p * W + ln : (int64_t)ln * W + p;
    const float j0 = o_p[q], j1 = o_p[HW + q], j2 = o_p[2 * HW + q], j3 = o_p[3 * HW + q];
    const float j4 = o_p[4 * HW + q], j5 = o_p[5 * HW + q], w = e_p[q];
    form = form && j0 == 0.f && j1 == j5 && j3 == 0.f;
    big = big || (j5 - j2) * w > 80.f || (j5 - j4) * w > 80.f || j5 * w > 80.f;
  }
  return a.buf_ok && __ballot(!form) == 0 && __ballot(big) == 0;
}

// One scanline (pass `dir`, row or column `ln` of image n) of the stereo kernels: the P1/P2
// sweep, the fast step or the general step.  `lds_h` is this wave's own LDS window copy.
// BV: a vertical P1/P2 sweep reads the base plane a.base (the other paths read the cost and
// both horizontal messages; the two give the same h bit for bit).
template <int LPL, bool VERT, bool FULL, int FASTK, bool OUTS, int PF, bool BV = false>
__device__ __forceinline__ void fwd_scanline(const FwdArgs &a, int n, int dir, int ln, float (&lds_h)[2][kWave * LPL]) {
  constexpr int ND = (3 + LPL - 1) / LPL;  // neighbour lanes reached by a +-3 window

  const int H = a.H, W = a.W, C = a.C, K = a.K;
  const int len = VERT ? H : W;
  const int nsteps = len - 1;
  const int rslot = dir == RIGHT ? LEFT : dir == LEFT ? RIGHT : dir == DOWN ? UP : DOWN;
  const bool fwd = (dir == RIGHT || dir == DOWN);
  const int lane = threadIdx.x & (kWave - 1);
  const int c0 = lane * LPL;
  const int delta = a.delta;
  const int nj = a.Nj > 1 ? n : 0;
  const int64_t HW = (int64_t)H * W;
  // in-plane pixel index of step s: q0 + s * dq
  const int64_t dq = VERT ? (fwd ? (int64_t)W : -(int64_t)W) : (fwd ? 1 : -1);
  const int64_t q0 = VERT ? (int64_t)(fwd ? 0 : H - 1) * W + ln : (int64_t)ln * W + (fwd ? 0 : W - 1);
  const float *cost_p = a.cost + (int64_t)n * HW * C;
  const float *x1_p = (VERT ? a.msg_out : a.msg_in) + ((int64_t)n * 4 + (VERT ? RIGHT : UP)) * HW * C;
  const float *x2_p = (VERT ? a.msg_out : a.msg_in) + ((int64_t)n * 4 + (VERT ? LEFT : DOWN)) * HW * C;
  float *out_p = a.msg_out + ((int64_t)n * 4 + rslot) * HW * C;
  uint8_t *arg_p = a.argmin ? a.argmin + ((int64_t)n * 4 + rslot) * HW * C : nullptr;
  const int64_t rs_base = ((int64_t)n * 4 + rslot) * HW;
  // lane j < K streams jump entry j of (dir, pixel), lane >= K the edge weight
  const float *par_p = lane < K ? a.jump + (((int64_t)nj * 4 + dir) * K + lane) * HW
                                : a.edge + ((int64_t)n * 4 + dir) * HW;
  // per-lane byte offsets of the label row (labels past C read label C-1)
  uint32_t voff[LPL];
  bool valid[LPL];
#pragma unroll
  for (int k = 0; k < LPL; ++k) {
    valid[k] = FULL || c0 + k < C;
    voff[k] = (uint32_t)(FULL ? c0 + k : min(c0 + k, C - 1)) * 4u;
  }
  auto ld_row = [&](const float *row, float *d) {
    if (FULL) {
      ld_lanes<LPL>(row, voff[0], d);
    } else {
#pragma unroll
      for (int k = 0; k < LPL; ++k) d[k] = *reinterpret_cast<const float *>(reinterpret_cast<const char *>(row) + voff[k]);
    }
  };

  if constexpr (FASTK > 0) {
    constexpr int KK = FASTK;  // jump length of the fast step (K = a.K at launch)
    bool off = false;  // any non-zero window offset on this scanline (jump entry 0)
    const float *o_p = a.jump + ((int64_t)nj * 4 + dir) * K * HW;
    if constexpr (FASTK == 6) {
      if (p1p2_scanline<VERT>(a, n, dir, ln, lane)) {
        fwd_sweep_p1p2<LPL, VERT, FULL, OUTS, PF, BV>(a, n, dir, rslot, nj, q0, dq, nsteps, lane, c0, valid);
        return;
      }
    }
    for (int p = lane; p < len; p += kWave) off = off || o_p[VERT ? (int64_t)p * W + ln : (int64_t)ln * W + p] != 0.f;
    if (__ballot(off) == 0) {
      float chain[LPL];
      ld_row(a.msg_in + (rs_base + q0) * C, chain);
#pragma unroll
      for (int k = 0; k < LPL; ++k)
        if (valid[k]) {
          out_p[q0 * C + c0 + k] = chain[k];
          if (OUTS) arg_p[q0 * C + c0 + k] = 0;
        }
      if (OUTS) {
        a.scale[rs_base + q0] = 0.f;  // every lane stores the same value
        a.gargmin[rs_base + q0] = 0;
      }

      float pp[PF];
      float pc[PF][LPL], q1[PF][LPL], q2[PF][LPL];
      auto load_ring = [&](int s, float *dc, float *d1, float *d2, float &dpar) {
        const int64_t q = q0 + (int64_t)s * dq;
        ld_row(cost_p + q * C, dc);
        ld_row(x1_p + q * C, d1);
        ld_row(x2_p + q * C, d2);
        dpar = par_p[q];
      };
#pragma unroll
      for (int i = 0; i < PF; ++i) {
        load_ring(i < nsteps ? i : (nsteps > 0 ? nsteps - 1 : 0), pc[i], q1[i], q2[i], pp[i]);
        __builtin_amdgcn_sched_barrier(0);
      }
      float Lprev = 0.f;
      auto step = [&](int i, int s) {
        // Every use of ring slot i is made to depend on the chain of the previous step: the
        // scheduler takes loads issued in an earlier loop iteration as complete and would
        // otherwise hoist their uses (and the waits for them) to the top of the loop body.
        float h[LPL];
#pragma unroll
        for (int k = 0; k < LPL; ++k) {
          float c = pc[i][k];
          if (VERT) asm("" : "+v"(c) : "v"(chain[k]));  // the vertical order adds the chain last
          const float v = VERT ? ((c + q1[i][k]) + q2[i][k]) + chain[k] : ((c + chain[k]) + q1[i][k]) + q2[i][k];
          h[k] = valid[k] ? v : __builtin_inff();
        }
        float pv = pp[i];
        asm("" : "+v"(pv) : "v"(chain[0]));  // no instruction: orders pv after the chain
        const float w = readlane_f(pv, KK);
        const float JK = readlane_f(pv, KK - 1);
        // global minimum
        int hk = float_key(h[0]);
#pragma unroll
        for (int k = 1; k < LPL; ++k) hk = min(hk, float_key(h[k]));
        const float Hs = key_to_float(wave_min_i(hk));
        // window minimum over offsets -3..3, last minimal candidate in offset order
        float Lh[ND][LPL], Rh[ND][LPL];
#pragma unroll
        for (int k = 0; k < LPL; ++k) { Lh[0][k] = shr1_inf(h[k]); Rh[0][k] = shl1_inf(h[k]); }
#pragma unroll
        for (int d = 1; d < ND; ++d)
#pragma unroll
          for (int k = 0; k < LPL; ++k) { Lh[d][k] = shr1_inf(Lh[d - 1][k]); Rh[d][k] = shl1_inf(Rh[d - 1][k]); }
        // candidates, their minimum (min3 tree), then the last offset attaining it
        float cand[7][LPL];
#pragma unroll
        for (int kk = -3; kk <= 3; ++kk) {
          const float fk = readlane_f(pv, min(max(KK / 2 - kk, 1), KK - 1));
#pragma unroll
          for (int k = 0; k < LPL; ++k) {
            const int e = k + kk;
            float hv;
            if (e >= 0 && e < LPL) {
              hv = h[e >= 0 && e < LPL ? e : 0];
            } else if (e < 0) {
              const int j = -e, d = (j + LPL - 1) / LPL, slot = d * LPL - j;
              hv = Lh[d - 1][slot];
            } else {
              const int j = e - LPL + 1, d = (j + LPL - 1) / LPL, slot = j - 1 - (d - 1) * LPL;
              hv = Rh[d - 1][slot];
            }
            cand[kk + 3][k] = fmaf(fk, w, hv);
          }
        }
        float mw[LPL];
#pragma unroll
        for (int k = 0; k < LPL; ++k)
          mw[k] = fminf(fminf(fminf(cand[0][k], cand[1][k]), fminf(cand[2][k], cand[3][k])),
                        fminf(fminf(cand[4][k], cand[5][k]), cand[6][k]));
        const float gv = fmaf(JK, w, Hs);
        // msg = min(window, gv); its winner is the last candidate equal to it, or the
        // truncation term when none is (a window value equal to gv wins, as in the reference)
        int arg[LPL];
#pragma unroll
        for (int k = 0; k < LPL; ++k) {
          const float msg = fminf(mw[k], gv);
          chain[k] = msg - gv;  // u: the next pixel's chain term
          arg[k] = kArgTrunc;
          if (OUTS) {
#pragma unroll
            for (int kk = 0; kk < 7; ++kk) arg[k] = cand[kk][k] == msg ? kk - 3 : arg[k];
          }
        }
        // ---- tail: label of the minimum, normalisation, stores
        // label of the minimum: the single lane attaining it by ballot; ties (rare) take the
        // largest tie key, the reference's tree order
        int lstar = 0;
        if (OUTS) {
          uint64_t mk[LPL];
          int cnt = 0;
#pragma unroll
          for (int k = 0; k < LPL; ++k) {
            mk[k] = __ballot(valid[k] && h[k] == Hs);
            cnt += __popcll(mk[k]);
          }
#pragma unroll
          for (int k = 0; k < LPL; ++k)
            if (mk[k]) lstar = (__ffsll((long long)mk[k]) - 1) * LPL + k;
          if (cnt > 1) {
            int tk = -1;
#pragma unroll
            for (int k = 0; k < LPL; ++k) tk = max(tk, (valid[k] && h[k] == Hs) ? tie_key(c0 + k, a.bits) : -1);
            lstar = tie_key(wave_max_i(tk), a.bits);
          }
        }
        // shift of the log-sum-exp: 0 in the u frame (gv) whenever every u >= (Fmin - JK) w
        // is >= -80, i.e. no term can underflow; the wave maximum of u otherwise (rare)
        float umax = 0.f;
        if (__ballot(lane >= 1 && lane < KK && (JK - pv) * w > 80.f) != 0) {
          int mk = float_key(valid[0] ? chain[0] : -__builtin_inff());
#pragma unroll
          for (int k = 1; k < LPL; ++k) mk = max(mk, float_key(valid[k] ? chain[k] : -__builtin_inff()));
          umax = key_to_float(wave_max_i(mk));
        }
        float se = 0.f;
#pragma unroll
        for (int k = 0; k < LPL; ++k) se += valid[k] ? fast_exp(chain[k] - umax) : 0.f;
        se = wave_sum_f(se);
        const float Lc = umax + fast_log(se);
        const int64_t qq = q0 + (int64_t)(s + 1) * dq;
        float outv[LPL];
#pragma unroll
        for (int k = 0; k < LPL; ++k) outv[k] = chain[k] - Lc;
        if (FULL) {
          st_lanes<LPL>(out_p + qq * C, voff[0], outv);
        } else {
#pragma unroll
          for (int k = 0; k < LPL; ++k)
            if (valid[k]) out_p[qq * C + c0 + k] = outv[k];
        }
        if (OUTS) {
#pragma unroll
          for (int k = 0; k < LPL; ++k)
            if (valid[k]) arg_p[qq * C + c0 + k] = (uint8_t)(arg[k] == kArgTrunc ? kFromTrunc : c0 + k + arg[k]);
          a.scale[rs_base + qq] = fast_exp(gv + Lc - Lprev);  // same value from every lane
          a.gargmin[rs_base + qq] = (uint8_t)lstar;
        }
        Lprev = Lc;
      };
      int s0 = 0;
      for (; s0 + PF <= nsteps; s0 += PF) {
#pragma unroll
        for (int i = 0; i < PF; ++i) {
          step(i, s0 + i);
          const int sn = s0 + i + PF;
          load_ring(sn < nsteps ? sn : nsteps - 1, pc[i], q1[i], q2[i], pp[i]);
        }
      }
#pragma unroll
      for (int i = 0; i < PF; ++i)
        if (s0 + i < nsteps) step(i, s0 + i);
      return;
    }
  }

  float chain[LPL];
  {
    const float *src = a.msg_in + (rs_base + q0) * C;
    ld_row(src, chain);
#pragma unroll
    for (int k = 0; k < LPL; ++k)
      if (valid[k]) {
        out_p[q0 * C + c0 + k] = chain[k];
        if (arg_p) arg_p[q0 * C + c0 + k] = 0;
      }
    if (lane == 0) {
      if (a.scale) a.scale[rs_base + q0] = 0.f;
      if (a.gargmin) a.gargmin[rs_base + q0] = 0;
    }
  }

  float pc[PF][LPL], p1[PF][LPL], p2[PF][LPL], pp[PF];
  auto load_step = [&](int s, float *dc, float *d1, float *d2, float &dp) {
    const int64_t q = q0 + (int64_t)s * dq;
    ld_row(cost_p + q * C, dc);
    ld_row(x1_p + q * C, d1);
    ld_row(x2_p + q * C, d2);
    dp = par_p[q];
  };
#pragma unroll
  for (int i = 0; i < PF; ++i) {
    load_step(i < nsteps ? i : (nsteps > 0 ? nsteps - 1 : 0), pc[i], p1[i], p2[i], pp[i]);
    // keep the slots in issue order: the loop's wait counts assume slot 0 is the oldest
    __builtin_amdgcn_sched_barrier(0);
  }

  for (int s0 = 0; s0 < nsteps; s0 += PF) {
#pragma unroll
    for (int i = 0; i < PF; ++i) {
      const int s = s0 + i;
      if (s < nsteps) {
        float h[LPL];
#pragma unroll
        for (int k = 0; k < LPL; ++k) {
          const float v = VERT ? ((pc[i][k] + p1[i][k]) + p2[i][k]) + chain[k]
                               : ((pc[i][k] + chain[k]) + p1[i][k]) + p2[i][k];
          h[k] = valid[k] ? v : __builtin_inff();
        }
        const float pv = pp[i];
        const float w = readlane_f(pv, K);
        const float o = readlane_f(pv, 0);
        const float JK = readlane_f(pv, K - 1);
        const float Jmid = readlane_f(pv, K / 2);

        // global minimum, then its label: the single minimal lane by ballot, the reference's
        // tree tie order otherwise
        int hk = float_key(h[0]);
#pragma unroll
        for (int k = 1; k < LPL; ++k) hk = min(hk, float_key(h[k]));
        const float Hs = key_to_float(wave_min_i(hk));
        int lstar = 0;
        {
          uint64_t mk[LPL];
          int cnt = 0;
#pragma unroll
          for (int k = 0; k < LPL; ++k) {
            mk[k] = __ballot(valid[k] && h[k] == Hs);
            cnt += __popcll(mk[k]);
          }
          if (cnt == 1) {
#pragma unroll
            for (int k = 0; k < LPL; ++k)
              if (mk[k]) lstar = (__ffsll((long long)mk[k]) - 1) * LPL + k;
          } else {
            float bv = __builtin_inff();
            int bk = -1;
#pragma unroll
            for (int k = 0; k < LPL; ++k)
              if (valid[k]) argmin_step(bv, bk, h[k], tie_key(c0 + k, a.bits));
            seg_argmin<64>(bv, bk);
            lstar = tie_key(bk, a.bits);
          }
        }
        const float gv = fmaf(JK, w, Hs);
        float msg[LPL];
        int arg[LPL];
#pragma unroll
        for (int k = 0; k < LPL; ++k) { msg[k] = gv; arg[k] = kArgTrunc; }

        if (o == 0.f && delta <= 3) {
          float Lh[ND][LPL], Rh[ND][LPL];
#pragma unroll
          for (int k = 0; k < LPL; ++k) { Lh[0][k] = shr1_inf(h[k]); Rh[0][k] = shl1_inf(h[k]); }
#pragma unroll
          for (int d = 1; d < ND; ++d)
#pragma unroll
            for (int k = 0; k < LPL; ++k) { Lh[d][k] = shr1_inf(Lh[d - 1][k]); Rh[d][k] = shl1_inf(Rh[d - 1][k]); }
#pragma unroll
          for (int kk = -3; kk <= 3; ++kk) {
            if (kk < -delta || kk > delta) continue;
            const float fk = readlane_f(pv, min(max(K / 2 - kk, 1), K - 1));
#pragma unroll
            for (int k = 0; k < LPL; ++k) {
              const int e = k + kk;  // label offset from this lane's first label
              float hv;
              if (e >= 0 && e < LPL) {
                hv = h[e >= 0 && e < LPL ? e : 0];
              } else if (e < 0) {
                const int j = -e, d = (j + LPL - 1) / LPL, slot = d * LPL - j;
                hv = Lh[d - 1][slot];
              } else {
                const int j = e - LPL + 1, d = (j + LPL - 1) / LPL, slot = j - 1 - (d - 1) * LPL;
                hv = Rh[d - 1][slot];
              }
              const float nv = fmaf(fk, w, hv);
              if (nv <= msg[k]) { msg[k] = nv; arg[k] = kk; }
            }
          }
        } else {
          // offset window or delta > 3: candidates from an LDS copy of h
          const int buf = s & 1;
#pragma unroll
          for (int k = 0; k < LPL; ++k) lds_h[buf][c0 + k] = h[k];
          wave_lds_sync();  // this wave's own window copy
          const int k_lo = (int)floorf(o - 0.5f) - delta;
          const int k_hi = (int)ceilf(o + 0.5f) + delta;
          int st[LPL], sp[LPL];
#pragma unroll
          for (int k = 0; k < LPL; ++k) {
            const int c = c0 + k;
            st[k] = sat_u16(roundf((float)(c - delta + 1) + o - 1.0f));
            sp[k] = min(sat_u16(roundf((float)(c + delta) + o + 1.0f)), C);
          }
          for (int kk = k_lo; kk <= k_hi; ++kk) {
            const float fk = readlane_f(pv, min(max(K / 2 - kk, 1), K - 1));
#pragma unroll
            for (int k = 0; k < LPL; ++k) {
              const int l = c0 + k + kk;
              if (l >= st[k] && l < sp[k]) {
                const float nv = fmaf(fk, w, lds_h[buf][l]);
                if (nv <= msg[k]) { msg[k] = nv; arg[k] = kk; }
              }
            }
          }
        }

        // normalisation shifted by gv >= every message; the label l* term keeps the sum away
        // from underflow unless (JK - J[K/2]) * w is huge, where the max-shift is used instead
        float shift = gv;
        if (!((JK - Jmid) * w <= 80.f)) {
          int mx = float_key(valid[0] ? msg[0] : -__builtin_inff());
#pragma unroll
          for (int k = 1; k < LPL; ++k) mx = max(mx, float_key(valid[k] ? msg[k] : -__builtin_inff()));
          shift = key_to_float(wave_max_i(mx));
        }
        float se = 0.f;
#pragma unroll
        for (int k = 0; k < LPL; ++k) se += valid[k] ? fast_exp(msg[k] - shift) : 0.f;
        se = wave_sum_f(se);
        const float lse = shift + fast_log(se);

        const int64_t qq = q0 + (int64_t)(s + 1) * dq;  // receiver
#pragma unroll
        for (int k = 0; k < LPL; ++k) chain[k] = msg[k] - lse;
        if (FULL) {
          st_lanes<LPL>(out_p + qq * C, voff[0], chain);
        } else {
#pragma unroll
          for (int k = 0; k < LPL; ++k)
            if (valid[k]) out_p[qq * C + c0 + k] = chain[k];
        }
        if (arg_p) {
#pragma unroll
          for (int k = 0; k < LPL; ++k)
            if (valid[k]) arg_p[qq * C + c0 + k] = (uint8_t)(arg[k] == kArgTrunc ? kFromTrunc : c0 + k + arg[k]);
        }
        if (lane == 0) {
          if (a.scale) a.scale[rs_base + qq] = fast_exp(lse);
          if (a.gargmin) a.gargmin[rs_base + qq] = (uint8_t)lstar;
        }
      }
      // refill ring slot i after its last use
      load_step(s + PF < nsteps ? s + PF : nsteps - 1, pc[i], p1[i], p2[i], pp[i]);
    }
  }
}

template <int LPL, bool VERT, bool FULL, int FASTK, bool OUTS, int PF, bool BV = false>
__global__ __launch_bounds__(64) void lbp_fwd_stereo_w64(FwdArgs a) {
  __shared__ float lds_h[2][kWave * LPL];
  const int L = VERT ? a.W : a.H;
  // columns of a vertical launch go to the XCDs in contiguous ranges (shared cache lines)
  const int bi = VERT ? xcd_contiguous_block((int)blockIdx.x, (int)gridDim.x) : (int)blockIdx.x;
  const int g = bi / L;
  const int dir = VERT ? ((g & 1) ? UP : DOWN) : ((g & 1) ? LEFT : RIGHT);
  fwd_scanline<LPL, VERT, FULL, FASTK, OUTS, PF, BV>(a, g >> 1, dir, bi - g * L, lds_h);
}

// Horizontal phase that also writes the base plane (cost + M_R) + M_L for the vertical phase
// (FwdArgs::base), with the fast step's launch conditions (delta 3, jump length 6).  A
// workgroup holds two rows of one image, four waves (the RIGHT and LEFT passes of each row; one
// wave per SIMD); with an odd row count the last workgroup's second row repeats row H - 1,
// whose stores then write the same values twice.  When all four scanlines have the P1/P2 form
// the passes run paired (the second visitor of a pixel writes its base entry, see
// fwd_sweep_p1p2); otherwise each wave runs its scanline alone.  After a final barrier the
// base entries not yet written (every entry when unpaired) are formed from the stored
// messages; the two waves of a row take alternate pixels.
template <int LPL, bool FULL, bool OUTS, int PF>
__global__ __launch_bounds__(4 * kWave) void lbp_fwd_stereo_pair(FwdArgs a) {
  __shared__ float lds_h_all[4][2][kWave * LPL];
  __shared__ int ok_all[4];
  const int wq = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int lane = threadIdx.x & (kWave - 1);
  const int H = a.H, W = a.W, C = a.C;
  const int H2 = (H + 1) / 2;
  const int n = (int)blockIdx.x / H2;
  const int y = min(2 * ((int)blockIdx.x - n * H2) + (wq >> 1), H - 1);
  const int dir = (wq & 1) ? LEFT : RIGHT;
  const int nsteps = W - 1;
  const bool ok = p1p2_scanline<false>(a, n, dir, y, lane);
  if (lane == 0) ok_all[wq] = ok ? 1 : 0;
  __syncthreads();
  const bool paired = ok_all[0] && ok_all[1] && ok_all[2] && ok_all[3];
  // second visits from step Tb on: the other pass reaches pixel x at step W - 2 - x, so every
  // receiver of a step >= Tb >= (W - 2) / 2 was reached by it before the barrier
  const int Tb = PF * (((W - 1) / 2 + PF - 1) / PF);
  int fix_lo = 0, fix_hi = W - 1;
  if (paired) {
    const int c0 = lane * LPL;
    bool valid[LPL];
#pragma unroll
    for (int k = 0; k < LPL; ++k) valid[k] = FULL || c0 + k < C;
    const int rslot = dir == RIGHT ? LEFT : RIGHT;
    const int64_t dq = dir == RIGHT ? 1 : -1;
    const int64_t q0 = (int64_t)y * W + (dir == RIGHT ? 0 : W - 1);
    fwd_sweep_p1p2<LPL, false, FULL, OUTS, PF, false, true>(a, n, dir, rslot, a.Nj > 1 ? n : 0, q0, dq, nsteps,
                                                            lane, c0, valid, Tb);
    if (Tb + PF <= nsteps - 1) {
      fix_lo = W - 1 - Tb;
      fix_hi = Tb;
    }
  } else {
    fwd_scanline<LPL, false, FULL, 6, OUTS, PF>(a, n, dir, y, lds_h_all[wq]);
  }
  __syncthreads();  // every message of both passes is written
  const int64_t HW = (int64_t)H * W, HWC = HW * C;
  const float *cp = a.cost + (int64_t)n * HWC;
  const float *mr = a.msg_out + ((int64_t)n * 4 + RIGHT) * HWC;
  const float *ml = a.msg_out + ((int64_t)n * 4 + LEFT) * HWC;
  float *bp = a.base + (int64_t)n * HWC;
  auto fix = [&](int x) {
    const int64_t o = ((int64_t)y * W + x) * C;
#pragma unroll
    for (int k = 0; k < LPL; ++k) {
      const int c = lane * LPL + k;
      if (FULL || c < C) bp[o + c] = (cp[o + c] + mr[o + c]) + ml[o + c];
    }
  };
  const int wv = wq & 1;
  int idx = 0;
  auto visit = [&](int x) {
    if ((idx++ & 1) == wv) fix(x);
  };
  if (fix_lo > 0) visit(0);
  for (int x = fix_lo; x <= fix_hi; ++x) visit(x);
  if (fix_hi < W - 1) visit(W - 1);
}

// prefetch depth (steps) of the stereo kernels' register ring, per phase
#ifndef BPL_FWD_PF_H
#define BPL_FWD_PF_H 8
#endif
#ifndef BPL_FWD_PF_V
#define BPL_FWD_PF_V 4
#endif
template <int LPL, bool FULL, int FASTK, bool OUTS>
void launch_stereo_w64_k(const FwdArgs &a, bool vert, hipStream_t st) {
  const dim3 grid((unsigned)(a.N * 2 * (vert ? a.W : a.H)));
  if (vert)
    hipLaunchKernelGGL((lbp_fwd_stereo_w64<LPL, true, FULL, FASTK, OUTS, BPL_FWD_PF_V>), grid, dim3(kWave), 0, st, a);
  else
    hipLaunchKernelGGL((lbp_fwd_stereo_w64<LPL, false, FULL, FASTK, OUTS, BPL_FWD_PF_H>), grid, dim3(kWave), 0, st, a);
}
// base-plane launches (a.base set by the host only with the fast step's conditions, K = 6):
// the horizontal phase writes the plane, the vertical P1/P2 sweeps read it
template <int LPL, bool FULL, bool OUTS>
void launch_stereo_base(const FwdArgs &a, bool vert, hipStream_t st) {
  if (vert) {
    const dim3 grid((unsigned)(a.N * 2 * a.W));
    hipLaunchKernelGGL((lbp_fwd_stereo_w64<LPL, true, FULL, 6, OUTS, BPL_FWD_PF_V, true>), grid, dim3(kWave), 0, st, a);
  } else {
    const dim3 grid((unsigned)(a.N * ((a.H + 1) / 2)));
    hipLaunchKernelGGL((lbp_fwd_stereo_pair<LPL, FULL, OUTS, BPL_FWD_PF_H>), grid, dim3(4 * kWave), 0, st, a);
  }
}
template <int LPL, bool FULL>
void launch_stereo_w64(const FwdArgs &a, bool vert, hipStream_t st) {
  // fast step: delta 3, the jump lengths of the stereo packings (K_in = 3 -> 6, K_in = 5 ->
  // 8), and argmins, labels and scales all requested (training) or none (inference)
  const bool all = a.argmin && a.gargmin && a.scale, none = !a.argmin && !a.gargmin && !a.scale;
  const bool fast = a.delta == 3 && (all || none);
  if (fast && a.K == 6 && a.base) {
    if (all) launch_stereo_base<LPL, FULL, true>(a, vert, st);
    else launch_stereo_base<LPL, FULL, false>(a, vert, st);
  } else if (fast && a.K == 6) {
    if (all) launch_stereo_w64_k<LPL, FULL, 6, true>(a, vert, st);
    else launch_stereo_w64_k<LPL, FULL, 6, false>(a, vert, st);
  } else if (fast && a.K == 8) {
    if (all) launch_stereo_w64_k<LPL, FULL, 8, true>(a, vert, st);
    else launch_stereo_w64_k<LPL, FULL, 8, false>(a, vert, st);
  } else {
    launch_stereo_w64_k<LPL, FULL, 0, false>(a, vert, st);
  }
}
template <int LPL>
void launch_stereo_w64_any(const FwdArgs &a, bool vert, hipStream_t st) {
  if (a.C == kWave * LPL) launch_stereo_w64<LPL, true>(a, vert, st);
  else launch_stereo_w64<LPL, false>(a, vert, st);
}

bool stereo_w64_ok(const FwdArgs &a) { return a.C > 16 && a.delta < a.C; }
// the semantic per-pixel scanline kernel: full min-plus branch, C <= 32, the messages of all
// images and a direction's C x C matrix plane below 2^31 bytes
bool semp_fast(const FwdArgs &a) {
  const int64_t hwc = (int64_t)a.H * a.W * a.C;
  return a.delta >= a.C && a.C <= 32 && (int64_t)a.N * 4 * hwc * 4 < (1LL << 31) &&
         hwc * a.C * 4 < (1LL << 31) && !generic_only();
}
// the semantic-global scanline kernel: full min-plus branch, C <= 32, [H,W,C] planes below 2^31 bytes
bool semg_fast(const FwdArgs &a) {
  return a.delta >= a.C && a.C <= 32 && (int64_t)a.H * a.W * a.C * 4 < (1LL << 31) && !generic_only();
}

void launch_stereo_w64_all(const FwdArgs &a, bool vert, hipStream_t st) {
  const int C = a.C;
  if (C <= 64) launch_stereo_w64_any<1>(a, vert, st);
  else if (C <= 128) launch_stereo_w64_any<2>(a, vert, st);
  else if (C <= 192) launch_stereo_w64_any<3>(a, vert, st);
  else launch_stereo_w64_any<4>(a, vert, st);
}

using FwdKernel = void (*)(FwdArgs);

template <int KIND, int SEG, int LPL>
void launch_phase(const FwdArgs &a, bool vert, hipStream_t st, size_t shm) {
  constexpr int PF = 4;
  constexpr int SPW = kWave / SEG;
  const int L = vert ? a.W : a.H;
  const int wpg = (L + SPW - 1) / SPW;
  const dim3 grid((unsigned)(a.N * 2 * wpg));
  if (vert)
    hipLaunchKernelGGL((lbp_fwd_phase<KIND, SEG, LPL, true, PF>), grid, dim3(kWave), shm, st, a);
  else
    hipLaunchKernelGGL((lbp_fwd_phase<KIND, SEG, LPL, false, PF>), grid, dim3(kWave), shm, st, a);
}

template <int KIND>
void launch_kind(const FwdArgs &a, bool vert, hipStream_t st, size_t shm) {
  const int C = a.C;
  if (C <= 16) launch_phase<KIND, 16, 1>(a, vert, st, shm);
  else if (C <= 32) launch_phase<KIND, 32, 1>(a, vert, st, shm);
  else if (C <= 64) launch_phase<KIND, 64, 1>(a, vert, st, shm);
  else if (C <= 128) launch_phase<KIND, 64, 2>(a, vert, st, shm);
  else if (C <= 192) launch_phase<KIND, 64, 3>(a, vert, st, shm);
  else launch_phase<KIND, 64, 4>(a, vert, st, shm);
}


}  // namespace
}  // namespace bpl

using namespace bpl;

// BPL_LBP_NO_BASE=1 turns the base plane of bpl_lbp_fwd_ws off (the plain path), for comparisons.
static bool base_off() {
  const char *e = getenv("BPL_LBP_NO_BASE");
  return e && e[0] == '1';
}

static int lbp_fwd_impl(int kind, const float *cost, const float *jump, int64_t jump_k, int64_t jump_n,
                        const float *edge, const float *msg_in, float *msg_out,
                        uint8_t *argmin, uint8_t *gargmin, float *scale,
                        int64_t N, int64_t H, int64_t W, int64_t C, int delta, void *ws, int64_t ws_bytes,
                        int flags, void *stream) {
  BPL_REQUIRE(kind >= 0 && kind <= 2, BPL_E_ARG, "bpl_lbp_fwd: unknown kind %d", kind);
  BPL_REQUIRE(N >= 1 && H >= 1 && W >= 1 && C >= 1, BPL_E_ARG,
              "bpl_lbp_fwd: sizes must be positive (N=%lld H=%lld W=%lld C=%lld)",
              (long long)N, (long long)H, (long long)W, (long long)C);
  BPL_REQUIRE(C <= 255, BPL_E_LIMIT, "bpl_lbp_fwd: C=%lld labels exceeds the supported 255", (long long)C);
  BPL_REQUIRE(N * 2 * (H > W ? H : W) < (1LL << 31) && H < (1 << 30) && W < (1 << 30), BPL_E_LIMIT,
              "bpl_lbp_fwd: grid too large");
  BPL_REQUIRE(delta >= 0, BPL_E_ARG, "bpl_lbp_fwd: delta must be >= 0");
  BPL_REQUIRE(cost && jump && edge && msg_in && msg_out, BPL_E_ARG, "bpl_lbp_fwd: null pointer");
  BPL_REQUIRE(!argmin || delta >= C || gargmin, BPL_E_ARG,
              "bpl_lbp_fwd: gargmin is required with argmin when delta < C");
  if (kind == BPL_LBP_STEREO) {
    BPL_REQUIRE(jump_k >= 2 && jump_k <= 15, BPL_E_LIMIT, "bpl_lbp_fwd: stereo jump K=%lld not in [2,15]",
                (long long)jump_k);
    BPL_REQUIRE(jump_n == 1 || jump_n == N, BPL_E_ARG, "bpl_lbp_fwd: jump batch %lld must be 1 or N",
                (long long)jump_n);
  }
  FwdArgs a;
  a.cost = cost; a.jump = jump; a.edge = edge; a.msg_in = msg_in; a.msg_out = msg_out;
  a.argmin = argmin; a.gargmin = gargmin; a.scale = scale;
  a.N = (int)N; a.H = (int)H; a.W = (int)W; a.C = (int)C;
  a.K = kind == BPL_LBP_STEREO ? (int)jump_k : 0;
  a.Nj = kind == BPL_LBP_STEREO ? (int)jump_n : 1;
  a.delta = delta; a.bits = tie_bits((int)C);
  // the P1/P2 sweep addresses an image's four message planes through one 32-bit offset range
  a.buf_ok = 4 * H * W * C * 4 < (1LL << 31) && jump_k * H * W * 4 < (1LL << 31) && !p1p2_off();
  a.mat_lds = (kind == BPL_LBP_SEMANTIC_GLOBAL && C <= 64) ? 1 : 0;
  // the base plane: stereo fast-step launches whose planes fit the buffer addressing, with a
  // workspace of one [N,H,W,C] plane that overlaps no tensor the call reads or writes
  a.base = nullptr;
  a.in_zero = (flags & BPL_FWD_MSG_IN_ZERO) ? 1 : 0;
  if (ws && kind == BPL_LBP_STEREO && jump_k == 6 && delta == 3 && stereo_w64_ok(a) && a.buf_ok &&
      !generic_only() && !base_off() && ws_bytes >= N * H * W * C * (int64_t)sizeof(float))
    a.base = static_cast<float *>(ws);
  const size_t shm = a.mat_lds ? (size_t)C * C * sizeof(float) : 0;
  hipStream_t st = as_stream(stream);
  for (int phase = 0; phase < 2; ++phase) {
    const bool vert = phase == 1;
    KernelTimer kt(vert ? BPL_KERNEL_LBP_FWD_V : BPL_KERNEL_LBP_FWD_H, st);
    switch (kind) {
      case BPL_LBP_STEREO:
        if (stereo_w64_ok(a) && !generic_only()) launch_stereo_w64_all(a, vert, st);
        else launch_kind<BPL_LBP_STEREO>(a, vert, st, shm);
        break;
      case BPL_LBP_SEMANTIC_GLOBAL:
        if (semg_fast(a) && a.C <= 20) launch_semg<20>(a, vert, st);
        else if (semg_fast(a)) launch_semg<32>(a, vert, st);
        else launch_kind<BPL_LBP_SEMANTIC_GLOBAL>(a, vert, st, shm);
        break;
      default:
        if (semp_fast(a) && a.C <= 20) launch_semp<20, 4>(a, vert, st);
        else if (semp_fast(a)) launch_semp<32, 2>(a, vert, st);
        else launch_kind<BPL_LBP_SEMANTIC_PIXEL>(a, vert, st, shm);
        break;
    }
    int rc = check_hip(hipGetLastError(), "bpl_lbp_fwd launch");
    if (rc) return rc;
  }
  return BPL_OK;
}

extern "C" int bpl_lbp_fwd(int kind, const float *cost, const float *jump, int64_t jump_k, int64_t jump_n,
                           const float *edge, const float *msg_in, float *msg_out,
                           uint8_t *argmin, uint8_t *gargmin, float *scale,
                           int64_t N, int64_t H, int64_t W, int64_t C, int delta, void *stream) {
  return lbp_fwd_impl(kind, cost, jump, jump_k, jump_n, edge, msg_in, msg_out, argmin, gargmin, scale, N, H, W, C,
                      delta, nullptr, 0, 0, stream);
}

extern "C" int64_t bpl_lbp_fwd_workspace_size(int kind, int64_t N, int64_t H, int64_t W, int64_t C) {
  if (kind != BPL_LBP_STEREO || N < 1 || H < 1 || W < 1 || C < 1) return 0;
  return N * H * W * C * (int64_t)sizeof(float);
}

extern "C" int bpl_lbp_fwd_ws(int kind, const float *cost, const float *jump, int64_t jump_k, int64_t jump_n,
                              const float *edge, const float *msg_in, float *msg_out,
                              uint8_t *argmin, uint8_t *gargmin, float *scale,
                              int64_t N, int64_t H, int64_t W, int64_t C, int delta, void *workspace,
                              int64_t workspace_bytes, int flags, void *stream) {
  BPL_REQUIRE((flags & ~BPL_FWD_MSG_IN_ZERO) == 0, BPL_E_ARG, "bpl_lbp_fwd_ws: unknown flags 0x%x", flags);
  BPL_REQUIRE(workspace_bytes >= 0, BPL_E_ARG, "bpl_lbp_fwd_ws: negative workspace size");
  if (workspace && N >= 1 && H >= 1 && W >= 1 && C >= 1) {
    // the workspace is written while the cost and both message tensors are read
    const auto apart = [](const void *p, int64_t pb, const void *q, int64_t qb) {
      const uintptr_t a0 = (uintptr_t)p, b0 = (uintptr_t)q;
      return !q || a0 + (uintptr_t)pb <= b0 || b0 + (uintptr_t)qb <= a0;
    };
    const int64_t plane = N * H * W * C * (int64_t)sizeof(float);
    BPL_REQUIRE(apart(workspace, workspace_bytes, cost, plane) && apart(workspace, workspace_bytes, msg_in, 4 * plane) &&
                    apart(workspace, workspace_bytes, msg_out, 4 * plane),
                BPL_E_ARG, "bpl_lbp_fwd_ws: the workspace overlaps the cost or a message tensor");
  }
  BPL_REQUIRE(!workspace || workspace_bytes >= bpl_lbp_fwd_workspace_size(kind, N, H, W, C), BPL_E_ARG,
              "bpl_lbp_fwd_ws: workspace of %lld bytes, %lld needed", (long long)workspace_bytes,
              (long long)bpl_lbp_fwd_workspace_size(kind, N, H, W, C));
  return lbp_fwd_impl(kind, cost, jump, jump_k, jump_n, edge, msg_in, msg_out, argmin, gargmin, scale, N, H, W, C,
                      delta, workspace, workspace_bytes, flags, stream);
}
