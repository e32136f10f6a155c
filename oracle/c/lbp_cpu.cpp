// CPU restatement of the reference's stereo LBP op and stereo SAD (TEST INFRASTRUCTURE).
//
// Used as the timed CPU baseline of bench.py and, in double precision, as a second
// oracle checked against the numpy restatement (tests/test_oracle_c.py).  Follows:
//   forward   ops/lbp_stereo/src/lbp_min_sum_kernel.cu:9-213 (per step: h = cost + three
//             incoming messages, global tree argmin, truncated window scan with '<=',
//             log-sum-exp normalisation), host pass order :461-491;
//   backward  exact adjoint of the forward (vertical adjoints, horizontal adjoints with the
//             vertical h-gradients as cross term, combine), the same decomposition the
//             reference backward uses (:567-610) with the input-message gradient added;
//   SAD       ops/sad/src/stereo_sad_kernel.cu:16-108.
// Scanlines of one phase run in parallel with OpenMP (both directions of a phase are
// independent); each scanline is a sequential recurrence.
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <vector>

namespace {

enum { LEFT = 0, RIGHT = 1, UP = 2, DOWN = 3 };

inline int64_t pix(int n, int y, int x, int H, int W) { return ((int64_t)n * H + y) * W + x; }
inline int64_t spix(int n, int s, int y, int x, int H, int W) { return (((int64_t)n * 4 + s) * H + y) * W + x; }
inline int rslot(int d) { return d == RIGHT ? LEFT : d == LEFT ? RIGHT : d == DOWN ? UP : DOWN; }

template <class T>
struct Problem {
  const T *cost, *jump, *edge;
  int K, Nj, N, H, W, C, delta;
};

// global argmin of the tree reduction: among equal minima the largest bit-reversed index
inline int tie_key(int l, int bits) {
  int r = 0;
  for (int b = 0; b < bits; ++b) r |= ((l >> b) & 1) << (bits - 1 - b);
  return r;
}

inline int round_half_away(double v) { return (int)(v >= 0 ? floor(v + 0.5) : ceil(v - 0.5)); }

// One message: h [C] -> msg_out [C] (normalised), arg [C], from-truncation flags, lstar.
template <class T>
void send(const T *h, const T *J, int K, T w, int C, int delta, int bits, T *out, int *arg, uint8_t *fg,
          int *lstar_out, T *scale_out) {
  std::vector<T> msg(C);
  if (delta >= C) {
    for (int c = 0; c < C; ++c) {
      T m = (T)1e15;
      int a = 0;
      for (int l = 0; l < C; ++l) {
        const int idx = std::min(std::max(K / 2 + c - l, 1), K - 1);
        const T nv = h[l] + J[idx] * w;
        if (nv <= m) { m = nv; a = l; }
      }
      msg[c] = m;
      arg[c] = a;
      fg[c] = 0;
    }
    *lstar_out = -1;
  } else {
    T hs = h[0];
    int ls = 0, lk = tie_key(0, bits);
    for (int l = 1; l < C; ++l) {
      const int k = tie_key(l, bits);
      if (h[l] < hs || (h[l] == hs && k > lk)) { hs = h[l]; ls = l; lk = k; }
    }
    const T o = J[0];
    const T gv = hs + J[K - 1] * w;
    for (int c = 0; c < C; ++c) {
      T m = gv;
      int a = ls;
      uint8_t g = 1;
      const int st = std::max(round_half_away((double)(c - delta + 1) + (double)o - 1.0), 0);
      const int sp = std::min(std::max(round_half_away((double)(c + delta) + (double)o + 1.0), 0), C);
      for (int l = st; l < sp; ++l) {
        const int idx = std::min(std::max(K / 2 + c - l, 1), K - 1);
        const T nv = h[l] + J[idx] * w;
        if (nv <= m) { m = nv; a = l; g = 0; }
      }
      msg[c] = m;
      arg[c] = a;
      fg[c] = g;
    }
    *lstar_out = ls;
  }
  T mx = msg[0];
  for (int c = 1; c < C; ++c) mx = std::max(mx, msg[c]);
  T s = 0;
  for (int c = 0; c < C; ++c) s += exp(msg[c] - mx);
  const T lse = mx + log(s);
  for (int c = 0; c < C; ++c) out[c] = msg[c] - lse;
  *scale_out = exp(lse);
}

template <class T>
void fwd_phase(const Problem<T> &P, const T *min_, T *mout, int32_t *argmin, uint8_t *fromg, bool vert) {
  const int H = P.H, W = P.W, C = P.C, K = P.K;
  const int L = vert ? W : H, len = vert ? H : W;
  int bits = 0;
  while ((1 << bits) < C) ++bits;
  const int64_t HW = (int64_t)H * W;
#pragma omp parallel
  {
    std::vector<T> h(C), J(K), out(C);
    std::vector<int> arg(C);
    std::vector<uint8_t> fg(C);
#pragma omp for schedule(dynamic, 1)
    for (int64_t job = 0; job < (int64_t)P.N * 2 * L; ++job) {
      const int n = (int)(job / (2 * L));
      const int dsel = (int)((job / L) % 2);
      const int line = (int)(job % L);
      const int d = vert ? (dsel ? UP : DOWN) : (dsel ? LEFT : RIGHT);
      const bool fwd = d == RIGHT || d == DOWN;
      const int r = rslot(d);
      const int nj = P.Nj > 1 ? n : 0;
      const T *cross = vert ? mout : min_;
      const int s1 = vert ? RIGHT : UP, s2 = vert ? LEFT : DOWN;
      auto yx = [&](int p, int &y, int &x) { if (vert) { y = p; x = line; } else { y = line; x = p; } };
      int y, x;
      yx(fwd ? 0 : len - 1, y, x);
      std::vector<T> chain(C);
      {
        const int64_t o = spix(n, r, y, x, H, W) * C;
        for (int c = 0; c < C; ++c) { chain[c] = min_[o + c]; mout[o + c] = chain[c]; argmin[o + c] = 0; fromg[o + c] = 0; }
      }
      for (int s = 0; s < len - 1; ++s) {
        yx(fwd ? s : len - 1 - s, y, x);
        const int64_t cp = pix(n, y, x, H, W) * C;
        const int64_t o1 = spix(n, s1, y, x, H, W) * C, o2 = spix(n, s2, y, x, H, W) * C;
        for (int c = 0; c < C; ++c)
          h[c] = vert ? ((P.cost[cp + c] + cross[o1 + c]) + cross[o2 + c]) + chain[c]
                      : ((P.cost[cp + c] + chain[c]) + cross[o1 + c]) + cross[o2 + c];
        for (int k = 0; k < K; ++k) J[k] = P.jump[(((int64_t)nj * 4 + d) * K + k) * HW + (int64_t)y * W + x];
        const T w = P.edge[spix(n, d, y, x, H, W)];
        int ls;
        T sc;
        send(h.data(), J.data(), K, w, C, P.delta, bits, out.data(), arg.data(), fg.data(), &ls, &sc);
        int yq = y, xq = x;
        if (vert) yq += fwd ? 1 : -1; else xq += fwd ? 1 : -1;
        const int64_t oq = spix(n, r, yq, xq, H, W) * C;
        for (int c = 0; c < C; ++c) {
          mout[oq + c] = out[c];
          argmin[oq + c] = arg[c];
          fromg[oq + c] = fg[c];
          chain[c] = out[c];
        }
      }
    }
  }
}

template <class T>
void bwd_phase(const Problem<T> &P, const T *G, const T *M, const int32_t *argmin, const uint8_t *fromg,
               T *X[4], T *grad_jump, T *grad_edge, bool vert) {
  const int H = P.H, W = P.W, C = P.C, K = P.K;
  const int L = vert ? W : H, len = vert ? H : W;
  const int64_t HW = (int64_t)H * W;
#pragma omp parallel
  {
    std::vector<T> acc(C), g(C), gh(C), J(K);
#pragma omp for schedule(dynamic, 1)
    for (int64_t job = 0; job < (int64_t)P.N * 2 * L; ++job) {
      const int n = (int)(job / (2 * L));
      const int dsel = (int)((job / L) % 2);
      const int line = (int)(job % L);
      const int d = vert ? (dsel ? UP : DOWN) : (dsel ? LEFT : RIGHT);
      const bool fwd = d == RIGHT || d == DOWN;
      const int r = rslot(d);
      const int nj = P.Nj > 1 ? n : 0;
      T *Xd = X[d];
      auto yx = [&](int p, int &y, int &x) { if (vert) { y = p; x = line; } else { y = line; x = p; } };
      int y, x;
      yx(fwd ? len - 1 : 0, y, x);
      for (int c = 0; c < C; ++c) Xd[pix(n, y, x, H, W) * C + c] = 0;
      grad_edge[spix(n, d, y, x, H, W)] = 0;
      for (int k = 0; k < K; ++k) grad_jump[(((int64_t)nj * 4 + d) * K + k) * HW + (int64_t)y * W + x] = 0;
      std::fill(acc.begin(), acc.end(), (T)0);
      for (int s = len - 2; s >= 0; --s) {
        yx(fwd ? s : len - 1 - s, y, x);
        int yq = y, xq = x;
        if (vert) yq += fwd ? 1 : -1; else xq += fwd ? 1 : -1;
        const int64_t oq = spix(n, r, yq, xq, H, W) * C;
        T ts = 0;
        for (int c = 0; c < C; ++c) ts += G[oq + c] + acc[c];
        const int64_t pq = pix(n, yq, xq, H, W) * C;
        for (int c = 0; c < C; ++c) {
          g[c] = (G[oq + c] + acc[c]) - exp(M[oq + c]) * ts;
          if (!vert) g[c] += X[DOWN][pq + c] + X[UP][pq + c];
        }
        for (int k = 0; k < K; ++k) J[k] = P.jump[(((int64_t)nj * 4 + d) * K + k) * HW + (int64_t)y * W + x];
        const T w = P.edge[spix(n, d, y, x, H, W)];
        std::fill(gh.begin(), gh.end(), (T)0);
        std::vector<T> bins(K, (T)0);
        for (int c = 0; c < C; ++c) {
          const int a = argmin[oq + c];
          gh[a] += g[c];
          const int b = fromg[oq + c] ? K - 1 : std::min(std::max(K / 2 + c - a, 1), K - 1);
          bins[b] += g[c];
        }
        T ge = 0;
        for (int k = 1; k < K; ++k) {
          ge += J[k] * bins[k];
          grad_jump[(((int64_t)nj * 4 + d) * K + k) * HW + (int64_t)y * W + x] = w * bins[k];
        }
        grad_jump[(((int64_t)nj * 4 + d) * K) * HW + (int64_t)y * W + x] = 0;
        grad_edge[spix(n, d, y, x, H, W)] = ge;
        const int64_t op = pix(n, y, x, H, W) * C;
        for (int c = 0; c < C; ++c) { acc[c] = gh[c]; Xd[op + c] = gh[c]; }
      }
    }
  }
}

template <class T>
void combine(const Problem<T> &P, T *X[4], const T *G, T *grad_cost, T *grad_msg) {
  const int H = P.H, W = P.W, C = P.C;
  const int64_t HWC = (int64_t)H * W * C;
#pragma omp parallel for
  for (int64_t i = 0; i < (int64_t)P.N * HWC; ++i) {
    const T xl = X[LEFT][i], xr = X[RIGHT][i], xu = X[UP][i], xd = X[DOWN][i];
    grad_cost[i] = ((xd + xu) + xr) + xl;
    const int64_t n = i / HWC, rr = i - n * HWC, base = n * 4 * HWC + rr;
    const int64_t yxp = rr / C;
    const int y = (int)(yxp / W), x = (int)(yxp % W);
    const T hr = xr + xl, sv = xd + xu;
    grad_msg[base + LEFT * HWC] = x == 0 ? (G[base + LEFT * HWC] + xr) + sv : 0;
    grad_msg[base + RIGHT * HWC] = x == W - 1 ? (G[base + RIGHT * HWC] + xl) + sv : 0;
    grad_msg[base + UP * HWC] = y == 0 ? (G[base + UP * HWC] + xd) + hr : hr;
    grad_msg[base + DOWN * HWC] = y == H - 1 ? (G[base + DOWN * HWC] + xu) + hr : hr;
  }
}

template <class T>
int lbp_fwd(const T *cost, const T *jump, int K, int Nj, const T *edge, const T *min_, T *mout, int32_t *argmin,
            uint8_t *fromg, int N, int H, int W, int C, int delta, int threads) {
  if (threads > 0) omp_set_num_threads(threads);
  Problem<T> P{cost, jump, edge, K, Nj, N, H, W, C, delta};
  fwd_phase(P, min_, mout, argmin, fromg, false);
  fwd_phase(P, min_, mout, argmin, fromg, true);
  return 0;
}

template <class T>
int lbp_bwd(const T *jump, int K, int Nj, const T *edge, const T *G, const T *M, const int32_t *argmin,
            const uint8_t *fromg, T *grad_cost, T *grad_jump, T *grad_edge, T *grad_msg, T *work, int N, int H,
            int W, int C, int delta, int threads) {
  if (threads > 0) omp_set_num_threads(threads);
  Problem<T> P{nullptr, jump, edge, K, Nj, N, H, W, C, delta};
  const int64_t plane = (int64_t)N * H * W * C;
  T *X[4] = {work, work + plane, work + 2 * plane, work + 3 * plane};
  bwd_phase(P, G, M, argmin, fromg, X, grad_jump, grad_edge, true);
  bwd_phase(P, G, M, argmin, fromg, X, grad_jump, grad_edge, false);
  combine(P, X, G, grad_cost, grad_msg);
  return 0;
}

template <class T>
void sad_fwd(const T *f0, const T *f1, T *cost, int N, int C, int H, int W, int mind, int D, int threads) {
  if (threads > 0) omp_set_num_threads(threads);
  const int64_t HW = (int64_t)H * W;
#pragma omp parallel for collapse(2)
  for (int n = 0; n < N; ++n)
    for (int y = 0; y < H; ++y)
      for (int x = 0; x < W; ++x)
        for (int i = 0; i < D; ++i) {
          const int d = mind + i;
          T r = 40;
          if (x - d >= 0 && x - d < W) {
            r = 0;
            for (int c = 0; c < C; ++c) {
              const int64_t b = ((int64_t)n * C + c) * HW + (int64_t)y * W;
              r += fabs(f0[b + x] - f1[b + x - d]);
            }
          }
          cost[pix(n, y, x, H, W) * D + i] = r;
        }
}

}  // namespace

extern "C" {
int cpu_lbp_fwd_f32(const float *cost, const float *jump, int K, int Nj, const float *edge, const float *min_,
                    float *mout, int32_t *argmin, uint8_t *fromg, int N, int H, int W, int C, int delta, int thr) {
  return lbp_fwd<float>(cost, jump, K, Nj, edge, min_, mout, argmin, fromg, N, H, W, C, delta, thr);
}
int cpu_lbp_fwd_f64(const double *cost, const double *jump, int K, int Nj, const double *edge, const double *min_,
                    double *mout, int32_t *argmin, uint8_t *fromg, int N, int H, int W, int C, int delta, int thr) {
  return lbp_fwd<double>(cost, jump, K, Nj, edge, min_, mout, argmin, fromg, N, H, W, C, delta, thr);
}
int cpu_lbp_bwd_f32(const float *jump, int K, int Nj, const float *edge, const float *G, const float *M,
                    const int32_t *argmin, const uint8_t *fromg, float *gc, float *gj, float *ge, float *gm,
                    float *work, int N, int H, int W, int C, int delta, int thr) {
  return lbp_bwd<float>(jump, K, Nj, edge, G, M, argmin, fromg, gc, gj, ge, gm, work, N, H, W, C, delta, thr);
}
int cpu_lbp_bwd_f64(const double *jump, int K, int Nj, const double *edge, const double *G, const double *M,
                    const int32_t *argmin, const uint8_t *fromg, double *gc, double *gj, double *ge, double *gm,
                    double *work, int N, int H, int W, int C, int delta, int thr) {
  return lbp_bwd<double>(jump, K, Nj, edge, G, M, argmin, fromg, gc, gj, ge, gm, work, N, H, W, C, delta, thr);
}
void cpu_sad_fwd_f32(const float *f0, const float *f1, float *cost, int N, int C, int H, int W, int mind, int D,
                     int thr) {
  sad_fwd<float>(f0, f1, cost, N, C, H, W, mind, D, thr);
}
int cpu_max_threads(void) { return omp_get_max_threads(); }
}
