"""FP64 restatement of the reference's min-sum LBP ops (TEST INFRASTRUCTURE ONLY).

Reference sources followed (paths relative to the reference repository):

* forward kernel   ops/lbp_stereo/src/lbp_min_sum_kernel.cu:9-213, host loop :424-500
* backward kernel  ops/lbp_stereo/src/lbp_min_sum_kernel.cu:215-417, host loop :513-621
* helpers          ops/lbp_stereo/src/util.cuh:40-132
* semantic global  ops/lbp_semantic_pw/src/lbp_min_sum_kernel.cu:9-417 (f = jump[0,d,l,c])
* semantic pixel   ops/lbp_semantic_pw_pixel/src/lbp_min_sum_kernel.cu:9-416 (f = jump[d,y,x,l,c])

Layouts (all as in the reference): cost [N,H,W,C]; edge [N,4,H,W];
messages [N,4,H,W,C] with slot s at pixel p holding the message p received
from its neighbour on side s; slot order LEFT=0, RIGHT=1, UP=2, DOWN=3
(util.cuh:7).  A pass named by its direction d writes slot RECV[d] at the
receiver q = p + delta_d.

Deliberate differences from the reference (documented per function):
* every batch element is processed (the reference forward only runs n=0);
* the log-sum-exp normalisation is max-shifted (identical away from fp32
  overflow/underflow, where the reference clamps the sum at 1e-45);
* gradients of batch-shared parameters are summed over the batch.

Pinned by tests/test_oracle_lbp.py: brute-force min-marginals on chains,
truncated == full branch on ordered pairwise terms, log-sum-exp invariant,
mirror symmetry, the tie-order rule against a literal tree simulation, FP64
finite differences for the exact adjoint, and mode R == mode E at one
iteration.
"""
import numpy as np

LEFT, RIGHT, UP, DOWN = 0, 1, 2, 3
PASS_ORDER = (RIGHT, LEFT, DOWN, UP)
# slot written at the receiver by a pass in direction d
RECV = {RIGHT: LEFT, LEFT: RIGHT, DOWN: UP, UP: DOWN}
# (dy, dx) from sender to receiver
STEP = {RIGHT: (0, 1), LEFT: (0, -1), DOWN: (1, 0), UP: (-1, 0)}
MAX_FLOAT = 1e15  # lbp_min_sum_kernel.cu:50


def _round_half_away(x):
    """CUDA round(): halves away from zero."""
    return np.where(x >= 0, np.floor(x + 0.5), np.ceil(x - 0.5))


def _sat_u16(x):
    """float -> unsigned short conversion as done by the CUDA compiler (saturating)."""
    return np.clip(x, 0, 65535).astype(np.int64)


# --------------------------------------------------------------------------------------
# pairwise term variants
# --------------------------------------------------------------------------------------
class StereoPairwise:
    """Per-pixel label-difference pairwise of ops/lbp_stereo.

    jump [Nj,4,K,H,W] (Nj in {1,N}); f(c,l) = jump[n,d,clamp(K//2 + c - l, 1, K-1),y,x]
    (util.cuh:103-119), truncation value jump[n,d,K-1] and window offset jump[n,d,0]
    (lbp_min_sum_kernel.cu:114-118).
    """

    kind = "stereo"

    def __init__(self, jump):
        self.jump = np.asarray(jump, dtype=np.float64)
        assert self.jump.ndim == 5 and self.jump.shape[1] == 4
        self.K = self.jump.shape[2]

    def _j(self, n, d, ys, xs):
        nj = n if self.jump.shape[0] > 1 else 0
        return self.jump[nj, d][:, ys, xs].T  # [P, K]

    def idx(self, c, l):
        return np.clip(self.K // 2 + (np.asarray(c) - np.asarray(l)), 1, self.K - 1)

    def f(self, n, d, ys, xs, c, l):
        """values [P, len(c)] for label pairs (c[i], l[i]) (c, l broadcastable to [P, Cc])."""
        J = self._j(n, d, ys, xs)
        idx = self.idx(c, l)
        idx = np.broadcast_to(idx, (J.shape[0],) + np.shape(idx)[-1:])
        return np.take_along_axis(J, idx, axis=1)

    def trunc(self, n, d, ys, xs):
        return self._j(n, d, ys, xs)[:, self.K - 1]

    def window(self, n, d, ys, xs, delta, C):
        """[start, stop) per pixel and receiver label (lbp_min_sum_kernel.cu:117-118)."""
        o = self._j(n, d, ys, xs)[:, 0:1]
        c = np.arange(C)[None, :]
        # round(c - (delta - 1) + offset - 1): integer part first, then the float offset
        start = _sat_u16(_round_half_away((c - (delta - 1)).astype(np.float64) + o - 1.0))
        stop = np.minimum(_sat_u16(_round_half_away((c + delta).astype(np.float64) + o + 1.0)), C)
        return start, stop

    def zeros_grad(self, N):
        return np.zeros_like(self.jump)

    def add_grad(self, g, n, d, ys, xs, c, l, val, trunc_mask=None):
        """g[n,d,idx,y,x] += val (val [P,C]); entries flagged in trunc_mask go to K-1.
        A batch-shared jump (leading dim 1) accumulates the whole batch."""
        nj = n if self.jump.shape[0] > 1 else 0
        idx = np.broadcast_to(self.idx(c, l), val.shape)
        if trunc_mask is not None:
            idx = np.where(trunc_mask, self.K - 1, idx)
        for k in range(1, self.K):
            s = np.where(idx == k, val, 0.0).sum(axis=1)
            g[nj, d, k, ys, xs] += s


class SemanticGlobalPairwise:
    """Global C x C pairwise of ops/lbp_semantic_pw: f(c,l) = jump[0,d,l,c] (kernel :104);
    truncation jump[0,d,0,C-1] (:62); window [max(c-delta+1,0), min(c+delta-1,C-1)] (:64-65)."""

    kind = "semantic_global"

    def __init__(self, jump):
        self.jump = np.asarray(jump, dtype=np.float64)
        assert self.jump.ndim == 4 and self.jump.shape[:2] == (1, 4)

    def f(self, n, d, ys, xs, c, l):
        M = self.jump[0, d]
        v = M[np.asarray(l), np.asarray(c)]
        return np.broadcast_to(v, (len(ys),) + np.shape(v)[-1:]).copy()

    def trunc(self, n, d, ys, xs):
        C = self.jump.shape[-1]
        return np.full(len(ys), self.jump[0, d, 0, C - 1])

    def window(self, n, d, ys, xs, delta, C):
        c = np.arange(C)[None, :]
        start = np.maximum(c - delta + 1, 0)
        stop = np.minimum(c + delta - 1, C - 1) + 1
        P = len(ys)
        return np.broadcast_to(start, (P, C)), np.broadcast_to(stop, (P, C))

    def zeros_grad(self, N):
        return np.zeros_like(self.jump)

    def add_grad(self, g, n, d, ys, xs, c, l, val, trunc_mask=None):
        C = self.jump.shape[-1]
        ll = np.broadcast_to(l, val.shape)
        cc = np.broadcast_to(c, val.shape)
        if trunc_mask is not None:
            ll = np.where(trunc_mask, 0, ll)
            cc = np.where(trunc_mask, C - 1, cc)
        np.add.at(g[0, d], (ll.ravel(), cc.ravel()), val.ravel())


class SemanticPixelPairwise:
    """Per-pixel C x C pairwise of ops/lbp_semantic_pw_pixel: f(c,l) = jump[d,y,x,l,c]
    (kernel :104,151); truncation jump[d,y,x,0,C-1] (:62)."""

    kind = "semantic_pixel"

    def __init__(self, jump):
        self.jump = np.asarray(jump, dtype=np.float64)
        assert self.jump.ndim == 5 and self.jump.shape[0] == 4

    def f(self, n, d, ys, xs, c, l):
        M = self.jump[d][ys, xs]  # [P, C, C]  (l, c)
        P = M.shape[0]
        ll = np.broadcast_to(l, (P,) + np.shape(l)[-1:])
        cc = np.broadcast_to(c, (P,) + np.shape(c)[-1:])
        return M[np.arange(P)[:, None], ll, cc]

    def trunc(self, n, d, ys, xs):
        C = self.jump.shape[-1]
        return self.jump[d][ys, xs][:, 0, C - 1]

    window = SemanticGlobalPairwise.window

    def zeros_grad(self, N):
        return np.zeros_like(self.jump)

    def add_grad(self, g, n, d, ys, xs, c, l, val, trunc_mask=None):
        C = self.jump.shape[-1]
        P = val.shape[0]
        ll = np.broadcast_to(l, val.shape)
        cc = np.broadcast_to(c, val.shape)
        if trunc_mask is not None:
            ll = np.where(trunc_mask, 0, ll)
            cc = np.where(trunc_mask, C - 1, cc)
        yy = np.broadcast_to(np.asarray(ys)[:, None], val.shape)
        xx = np.broadcast_to(np.asarray(xs)[:, None], val.shape)
        np.add.at(g[d], (yy.ravel(), xx.ravel(), ll.ravel(), cc.ravel()), val.ravel())


def make_pairwise(kind, jump):
    return {"stereo": StereoPairwise, "semantic_global": SemanticGlobalPairwise,
            "semantic_pixel": SemanticPixelPairwise}[kind](jump)


# --------------------------------------------------------------------------------------
# forward
# --------------------------------------------------------------------------------------
def tree_argmin(h):
    """Global (min, argmin) over labels by the reference's shared-memory tree reduction
    (lbp_min_sum_kernel.cu:121-142): block width B = nextpow2(C); at stride s the pair
    (t, t+s) with t+s < C keeps t+s when h[t+s] <= h[t].  h: [P, C]."""
    P, C = h.shape
    B = 1 << max(C - 1, 0).bit_length()
    v = np.full((P, B), MAX_FLOAT)
    v[:, :C] = h
    lab = np.zeros((P, B), dtype=np.int64)
    lab[:, :C] = np.arange(C)
    s = B // 2
    while s > 0:
        t = np.arange(s)
        t = t[t + s < C]
        take = v[:, t + s] <= v[:, t]
        v[:, t] = np.where(take, v[:, t + s], v[:, t])
        lab[:, t] = np.where(take, lab[:, t + s], lab[:, t])
        s //= 2
    return v[:, 0], lab[:, 0]


def send(h, pw, n, d, ys, xs, w, delta, want_margin=False):
    """One message computation for a scanline of senders (lbp_min_sum_kernel.cu:62-212).

    h [P,C] is cost + the three incoming messages other than the one from the receiver;
    w [P] the edge weights.  Returns (normalised message [P,C], argmin [P,C],
    scale [P] = sum_c exp(msg), from_global [P,C] = value came from the truncation term,
    global argmin [P] (-1 on the full branch)).
    """
    P, C = h.shape
    cs = np.arange(C)
    cands = []  # (value, label, is_global) for the margin diagnostics
    if delta >= C:
        # full branch (:88-111): running fminf over ascending labels, argmin = last attaining
        msg = np.full((P, C), MAX_FLOAT)
        arg = np.zeros((P, C), dtype=np.int64)
        for l in range(C):
            new = h[:, l:l + 1] + pw.f(n, d, ys, xs, cs, np.full(C, l)) * w[:, None]
            msg = np.minimum(msg, new)
            arg = np.where(msg == new, l, arg)
            if want_margin:
                cands.append((new, np.full((P, C), l), False))
        from_global = np.zeros((P, C), dtype=bool)
        lstar = np.full(P, -1, dtype=np.int64)
    else:
        # truncated branch (:112-164)
        Hs, lstar = tree_argmin(h)
        msg = np.broadcast_to((Hs + pw.trunc(n, d, ys, xs) * w)[:, None], (P, C)).copy()
        arg = np.broadcast_to(lstar[:, None], (P, C)).copy()
        from_global = np.ones((P, C), dtype=bool)
        if want_margin:
            cands.append((msg.copy(), arg.copy(), True))
            # other labels attaining (almost) the global min are alternative global winners
            Hs_sorted = np.sort(h, axis=1)
            second = Hs_sorted[:, 1] if C > 1 else np.full(P, np.inf)
            cands.append((np.broadcast_to((second + pw.trunc(n, d, ys, xs) * w)[:, None], (P, C)),
                          np.full((P, C), -2), True))
        start, stop = pw.window(n, d, ys, xs, delta, C)
        kmin = int((start - cs[None, :]).min())
        kmax = int((stop - 1 - cs[None, :]).max())
        for k in range(kmin, kmax + 1):
            l = cs + k
            valid = (l[None, :] >= start) & (l[None, :] < stop) & (l >= 0)[None, :] & (l < C)[None, :]
            lc = np.clip(l, 0, C - 1)
            new = h[:, lc] + pw.f(n, d, ys, xs, cs, lc) * w[:, None]
            take = valid & (new <= msg)
            msg = np.where(take, new, msg)
            arg = np.where(take, lc[None, :], arg)
            from_global &= ~take
            if want_margin:
                cands.append((np.where(valid, new, np.inf), np.broadcast_to(lc[None, :], (P, C)), False))
    # normalisation (:166-181), max-shifted
    m = msg.max(axis=1, keepdims=True)
    lse = m + np.log(np.exp(msg - m).sum(axis=1, keepdims=True))
    if want_margin:
        # gap to the best candidate with another label, and to a same-label candidate of the
        # other source unless the two are exactly equal (a structural tie that fp32 resolves
        # by the same '<=' rule, e.g. equal window and truncation entries)
        margin = np.full((P, C), np.inf)
        for v, lab, isg in cands:
            other_label = lab != arg
            other_source = (lab == arg) & (isg != from_global)
            gap = v - msg
            margin = np.minimum(margin, np.where(other_label, gap, np.inf))
            margin = np.minimum(margin, np.where(other_source & (gap != 0), np.abs(gap), np.inf))
        return msg - lse, arg, np.exp(lse[:, 0]), from_global, lstar, margin
    return msg - lse, arg, np.exp(lse[:, 0]), from_global, lstar


def _steps(d, H, W):
    """(sender line index) order of one pass (lbp_min_sum_kernel.cu:461-491)."""
    if d == RIGHT:
        return range(0, W - 1)
    if d == LEFT:
        return range(W - 1, 0, -1)
    if d == DOWN:
        return range(0, H - 1)
    return range(H - 1, 0, -1)


def _line(d, s, H, W):
    """sender pixels (ys, xs) and receivers (yq, xq) of step s of pass d."""
    if d in (LEFT, RIGHT):
        ys = np.arange(H)
        xs = np.full(H, s)
    else:
        ys = np.full(W, s)
        xs = np.arange(W)
    dy, dx = STEP[d]
    return ys, xs, ys + dy, xs + dx


def lbp_forward(cost, pw, edge, messages, delta, want_margin=False):
    """One LBP iteration: passes RIGHT, LEFT, DOWN, UP (lbp_min_sum_kernel.cu:424-500).

    Messages are updated in place in a copy of ``messages`` (the reference mutates its
    argument).  Returns (messages, argmin, scale, from_global, global_argmin) where
    argmin/from_global are [N,4,H,W,C], scale/global_argmin [N,4,H,W] (slot = receiver
    slot; unwritten entries 0 / -1).
    """
    cost = np.asarray(cost, dtype=np.float64)
    edge = np.asarray(edge, dtype=np.float64)
    M = np.array(messages, dtype=np.float64, copy=True)
    N, H, W, C = cost.shape
    argmin = np.zeros((N, 4, H, W, C), dtype=np.int64)
    scale = np.zeros((N, 4, H, W))
    from_global = np.zeros((N, 4, H, W, C), dtype=bool)
    gargmin = np.full((N, 4, H, W), -1, dtype=np.int64)
    margin = np.full((N, 4, H, W, C), np.inf)
    for n in range(N):
        for d in PASS_ORDER:
            r = RECV[d]
            for s in _steps(d, H, W):
                ys, xs, yq, xq = _line(d, s, H, W)
                h = cost[n, ys, xs, :].copy()
                for slot in (RIGHT, LEFT, UP, DOWN):  # summation order of :69-72
                    if slot != d:
                        h += M[n, slot, ys, xs, :]
                w = edge[n, d, ys, xs]
                res = send(h, pw, n, d, ys, xs, w, delta, want_margin)
                mn, a, S, fg, ls = res[:5]
                if want_margin:
                    margin[n, r, yq, xq, :] = res[5]
                M[n, r, yq, xq, :] = mn
                argmin[n, r, yq, xq, :] = a
                scale[n, r, yq, xq] = S
                from_global[n, r, yq, xq, :] = fg
                gargmin[n, r, yq, xq] = ls
    if want_margin:
        return M, argmin, scale, from_global, gargmin, margin
    return M, argmin, scale, from_global, gargmin


# --------------------------------------------------------------------------------------
# argmins of one fp32 iteration, replayed step by step from its own inputs
# --------------------------------------------------------------------------------------
U32 = 2.0 ** -24    # unit roundoff of fp32
ROUND_KAPPA = 6.0   # roundings bounded per candidate value (see local_argmins)


def _step_decisions(ops, fw_of, pw, n, d, ys, xs, w, delta):
    """FP64 decisions of one step from its fp32 operands, with an fp32 ambiguity mask.

    ops [P,C,4]: the four fp32 summands of h per sender label (cost, chain, two cross
    messages); fw_of(c, l) -> f(c,l) * w [P,C] for receiver labels c and sender labels l.
    Returns (label [P,C] of the winning sender (the global argmin where the truncation term
    won), from_global [P,C], clear [P,C]): ``clear`` is False where another candidate, or the
    global argmin's runner-up, lies within the fp32 rounding bound of the winner."""
    P, C, _ = ops.shape
    h = ops.sum(axis=2)
    A = np.abs(ops).sum(axis=2)
    bits = np.ascontiguousarray(ops.astype(np.float32)).view(np.uint32)   # operand identity
    rows = np.arange(P)[:, None]
    cs = np.arange(C)

    def same_ops(l1, l2):
        return (bits[rows, l1] == bits[rows, l2]).all(axis=-1)

    res = send(h, pw, n, d, ys, xs, w, delta)
    arg, fg, lstar = res[1], res[3], res[4]
    # candidates in the order of send(): (value, sender label, is_global, f*w)
    cands = []
    if delta >= C:
        for l in range(C):
            ll = np.full((P, C), l)
            fw = fw_of(cs[None, :], ll)
            cands.append((h[:, l:l + 1] + fw, ll, False, fw, np.ones((P, C), bool)))
    else:
        ls = np.broadcast_to(lstar[:, None], (P, C))
        ft = (pw.trunc(n, d, ys, xs) * w)[:, None]
        cands.append((np.broadcast_to(h[rows, ls] + ft, (P, C)), ls, True, np.broadcast_to(ft, (P, C)),
                      np.ones((P, C), bool)))
        start, stop = pw.window(n, d, ys, xs, delta, C)
        kmin = int((start - cs[None, :]).min())
        kmax = int((stop - 1 - cs[None, :]).max())
        for k in range(kmin, kmax + 1):
            l = cs + k
            valid = (l[None, :] >= start) & (l[None, :] < stop) & (l >= 0)[None, :] & (l < C)[None, :]
            lc = np.broadcast_to(np.clip(l, 0, C - 1)[None, :], (P, C))
            fw = fw_of(cs[None, :], lc)
            cands.append((h[rows, lc] + fw, lc, False, fw, valid))
    vals = np.stack([c[0] for c in cands])                  # [M,P,C]
    labs = np.stack([c[1] for c in cands])
    isg = np.array([c[2] for c in cands])[:, None, None]
    fws = np.stack([c[3] for c in cands])
    ok = np.stack([c[4] for c in cands])
    # frame allowance: an fp32 kernel may carry the chain shifted by up to log C plus the
    # largest pairwise value (a normalisation it has not applied yet)
    fmax = np.where(ok, np.abs(fws), 0.0).max(axis=(0, 2))[:, None]
    err_h = ROUND_KAPPA * U32 * (A + np.maximum(np.log(C), fmax))   # per sender label [P,C]
    wi = np.argmax((labs == arg[None]) & (isg == fg[None]) & ok, axis=0)   # the winning candidate
    take = lambda a: np.take_along_axis(a, wi[None], axis=0)[0]
    v_w, l_w, fw_w = take(vals), take(labs), take(fws)
    e_w = err_h[rows, l_w] + ROUND_KAPPA * U32 * np.abs(fw_w)
    clear = np.ones((P, C), dtype=bool)
    for m in range(len(cands)):
        other = ok[m] & ((labs[m] != arg) | (isg[m, 0, 0] != fg))
        e_m = err_h[rows, labs[m]] + ROUND_KAPPA * U32 * np.abs(fws[m])
        gap = np.abs(vals[m] - v_w)
        near = gap <= e_m + e_w
        exact = (vals[m] == v_w) & same_ops(labs[m], l_w) & (fws[m] == fw_w)
        clear &= ~(other & near & ~exact)
    if delta < C:
        # the label stored for truncation winners is the global argmin: its runner-up must be
        # outside the rounding bound too (or an identical-operand tie, resolved by the tree)
        ls = lstar[:, None]
        gap = np.abs(h - h[rows, ls])
        near = (gap <= err_h + err_h[rows, ls]) & (cs[None, :] != ls)
        exact = (h == h[rows, ls]) & same_ops(np.broadcast_to(cs[None, :], (P, C)), np.broadcast_to(ls, (P, C)))
        lab_ok = ~np.any(near & ~exact, axis=1)
        clear &= ~fg | lab_ok[:, None]
    label = np.where(fg, np.broadcast_to(lstar[:, None], (P, C)), arg)
    return label, fg, clear


def local_argmins(cost, pw, edge, msg_in, msg_out, delta):
    """Argmins of one LBP iteration recomputed in FP64 step by step from the fp32 inputs that
    a kernel actually used: for every pass and sender, h = cost + the sender's chain message
    (msg_out's slot of the pass, the value the pass wrote one step earlier, or the input at
    the first sender) + the two cross messages (msg_in's UP/DOWN slots for the horizontal
    passes, msg_out's LEFT/RIGHT slots for the vertical ones), exactly as the passes of
    lbp_forward read them (lbp_min_sum_kernel.cu:66-73, in-place updates in pass order).

    Returns (label [N,4,H,W,C], from_global [N,4,H,W,C], clear [N,4,H,W,C], written
    [N,4,H,W]).  label is the sender label the winner came from (the global argmin for
    truncation winners).  ``clear`` marks the receiver labels whose decision is fixed under
    fp32 rounding: every other candidate (another sender label, or the same label through
    the other source) differs from the winner by more than the sum of their rounding bounds,
    or is an exact tie of identical fp32 operands (decided by the same rule on every
    implementation).  A candidate value h_l + f w computed in fp32 from these inputs errs by
    at most ROUND_KAPPA * 2^-24 * (sum |operands of h_l| + max(log C, max|f w|) + |f w|):
    three additions and one fused multiply-add, the rounding of the stored chain message, and
    a chain carried in a shifted frame (msg - min instead of msg - lse: the two differ by the
    log-sum-exp of non-positive values, which lies in [-max|f w|, log C]; a constant shift of
    h changes no decision but scales the roundings).
    """
    cost = np.asarray(cost, dtype=np.float32)
    edge = np.asarray(edge, dtype=np.float64)
    Mi = np.asarray(msg_in, dtype=np.float32)
    Mo = np.asarray(msg_out, dtype=np.float32)
    N, H, W, C = cost.shape
    label = np.zeros((N, 4, H, W, C), dtype=np.int64)
    fgl = np.zeros((N, 4, H, W, C), dtype=bool)
    clear = np.ones((N, 4, H, W, C), dtype=bool)
    written = np.zeros((N, 4, H, W), dtype=bool)
    for n in range(N):
        for d in PASS_ORDER:
            r = RECV[d]
            horiz = d in (LEFT, RIGHT)
            cross = (UP, DOWN) if horiz else (LEFT, RIGHT)
            for s in _steps(d, H, W):
                ys, xs, yq, xq = _line(d, s, H, W)
                ops = np.stack([cost[n, ys, xs, :], Mo[n, r, ys, xs, :],
                                (Mi if horiz else Mo)[n, cross[0], ys, xs, :],
                                (Mi if horiz else Mo)[n, cross[1], ys, xs, :]], axis=2).astype(np.float64)
                w = edge[n, d, ys, xs]
                fw_of = lambda c, l: pw.f(n, d, ys, xs, np.broadcast_to(c, l.shape) if np.ndim(l) > 1 else c, l) \
                    * w[:, None]
                lab, fg, ok = _step_decisions(ops, fw_of, pw, n, d, ys, xs, w, delta)
                label[n, r, yq, xq] = lab
                fgl[n, r, yq, xq] = fg
                clear[n, r, yq, xq] = ok
                written[n, r, yq, xq] = True
    return label, fgl, clear, written


def lbp_forward_iters(cost, pw, edge, messages, delta, iters):
    """``iters`` iterations; returns the final messages and the per-iteration records
    (msg_in, msg_out, argmin, scale, from_global) needed by the backward passes."""
    recs = []
    M = np.asarray(messages, dtype=np.float64)
    for _ in range(iters):
        Mo, a, S, fg, _ls = lbp_forward(cost, pw, edge, M, delta)
        recs.append((M, Mo, a, S, fg))
        M = Mo
    return M, recs


# --------------------------------------------------------------------------------------
# backward, mode E: exact adjoint of lbp_forward
# --------------------------------------------------------------------------------------
def lbp_backward_exact(cost, pw, edge, msg_out, argmin, from_global, grad_out, delta=None, absolute=False):
    """Exact reverse-mode derivative of one ``lbp_forward`` iteration (argmins held fixed).

    Replays the forward steps in reverse.  For each written message m = msg - lse(msg):
    d msg = g - softmax(msg) * sum(g); the min routes d msg(c) to h(argmin(c)), and to
    the pairwise entry / edge weight that produced the winning value (the truncation entry
    when the value came from the global term).  h = cost + the three other incoming slots,
    so its gradient flows to the cost and to those slots' values at that time.

    Returns (grad_cost [N,H,W,C], grad_pairwise (shape of pw.zeros_grad(N)),
    grad_edge [N,4,H,W], grad_messages_in [N,4,H,W,C]).

    ``absolute=True`` returns the componentwise condition magnitudes instead: the same
    computation with every term replaced by its absolute value (|grad_out|, gw + p * sum(gw)
    for the normalisation Jacobian, |f| and |w| for the pairwise and edge terms), i.e. the sum
    of |term| over all the products that are added into each output.  An fp32 evaluation of
    the adjoint errs on an output by at most (number of roundings along its path) * 2^-24
    times this magnitude, so it is the scale against which a relative error of an output
    that results from cancellation is measured.
    """
    cost = np.asarray(cost, dtype=np.float64)
    edge = np.asarray(edge, dtype=np.float64)
    N, H, W, C = cost.shape
    gM = np.array(grad_out, dtype=np.float64, copy=True)
    if absolute:
        gM = np.abs(gM)
    g_cost = np.zeros_like(cost)
    g_edge = np.zeros((N, 4, H, W))
    g_pw = pw.zeros_grad(N)
    cs = np.arange(C)
    sgn = -1.0 if not absolute else 1.0
    mag = np.abs if absolute else (lambda v: v)
    for n in range(N):
        for d in reversed(PASS_ORDER):
            r = RECV[d]
            for s in reversed(list(_steps(d, H, W))):
                ys, xs, yq, xq = _line(d, s, H, W)
                gw = gM[n, r, yq, xq, :].copy()
                gM[n, r, yq, xq, :] = 0.0
                p = np.exp(msg_out[n, r, yq, xq, :])
                gmsg = gw + sgn * p * gw.sum(axis=1, keepdims=True)
                a = argmin[n, r, yq, xq, :]
                fg = from_global[n, r, yq, xq, :]
                P = len(ys)
                gh = np.zeros((P, C))
                np.add.at(gh, (np.repeat(np.arange(P), C), a.ravel()), gmsg.ravel())
                w = mag(edge[n, d, ys, xs])
                fv = mag(np.where(fg, pw.trunc(n, d, ys, xs)[:, None], pw.f(n, d, ys, xs, cs[None, :], a)))
                g_edge[n, d, ys, xs] += (fv * gmsg).sum(axis=1)
                pw.add_grad(g_pw, n, d, ys, xs, cs[None, :], a, w[:, None] * gmsg, trunc_mask=fg)
                g_cost[n, ys, xs, :] += gh
                for slot in range(4):
                    if slot != d:
                        gM[n, slot, ys, xs, :] += gh
    return g_cost, g_pw, g_edge, gM


def lbp_backward_exact_iters(cost, pw, edge, recs, grad_final, absolute=False):
    """Chain ``lbp_backward_exact`` over the iterations recorded by lbp_forward_iters
    (``absolute``: the condition magnitudes of the chained adjoint)."""
    G = np.asarray(grad_final, dtype=np.float64)
    gc = gp = ge = None
    for (Mi, Mo, a, S, fg) in reversed(recs):
        c_, p_, e_, G = lbp_backward_exact(cost, pw, edge, Mo, a, fg, G, absolute=absolute)
        gc = c_ if gc is None else gc + c_
        gp = p_ if gp is None else gp + p_
        ge = e_ if ge is None else ge + e_
    return gc, gp, ge, G


# --------------------------------------------------------------------------------------
# backward, mode R: the reference's backward kernel restated
# --------------------------------------------------------------------------------------
_BWD_SHIFT = {LEFT: (0, 1), RIGHT: (0, -1), DOWN: (-1, 0), UP: (1, 0)}   # :279-300
_BWD_GXY = {UP: DOWN, DOWN: UP, LEFT: RIGHT, RIGHT: LEFT}                  # :302-318


def lbp_backward_reference(cost, pw, edge, in_grad, messages, argmin, scale, absolute=False):
    """The reference's backward_minsum (lbp_min_sum_kernel.cu:215-417, host :513-621).

    Adjoint passes in the order of the host loop: slot DOWN (adjoint of the UP pass,
    senders y=1..H-1), slot UP (adjoint of DOWN, y=H-2..0), slot LEFT (adjoint of RIGHT,
    x=W-2..0), slot RIGHT (adjoint of LEFT, x=1..W-1).  Per receiver label c:
      hor  = normalised(in_grad) + normalised(acc_hor)          (:326-369)
      up   = saved[UP,q] + acc_up, down = saved[DOWN,q] + acc_down (horizontal only, :380-386)
    routed to the sender's argmin label for unary / pairwise / edge gradients (:391-402)
    and into the next accumulators (:404-412); saved[slot,p] = acc_hor (:416).
    The pairwise bin is always idx(c - argmin) and the edge term f(c, argmin) (util.cuh:111-132).
    grad_messages = in_grad (:534-536).  Batch-shared parameter gradients are summed over
    the batch and edge gradients are per batch element (the reference indexes them with 0).
    ``absolute=True``: the componentwise condition magnitudes (every term by its absolute
    value; see lbp_backward_exact).
    """
    cost = np.asarray(cost, dtype=np.float64)
    edge = np.asarray(edge, dtype=np.float64)
    in_grad = np.asarray(in_grad, dtype=np.float64)
    if absolute:
        in_grad = np.abs(in_grad)
    sgn = 1.0 if absolute else -1.0
    mag = np.abs if absolute else (lambda v: v)
    messages = np.asarray(messages, dtype=np.float64)
    scale = np.asarray(scale, dtype=np.float64)
    N, H, W, C = cost.shape
    g_unary = np.zeros_like(cost)
    g_edge = np.zeros((N, 4, H, W))
    g_pw = pw.zeros_grad(N)
    cs = np.arange(C)
    for n in range(N):
        saved = np.zeros((4, H, W, C))
        for bdir, order in ((DOWN, range(1, H)), (UP, range(H - 2, -1, -1)),
                            (LEFT, range(W - 2, -1, -1)), (RIGHT, range(1, W))):
            cross = bdir in (LEFT, RIGHT)
            L = H if cross else W
            acc = np.zeros((L, 3, C))
            gxy = _BWD_GXY[bdir]
            dy, dx = _BWD_SHIFT[bdir]
            for s in order:
                if cross:
                    ys = np.arange(H)
                    xs = np.full(H, s)
                else:
                    ys = np.full(W, s)
                    xs = np.arange(W)
                yq, xq = ys + dy, xs + dx
                w = mag(edge[n, gxy, ys, xs])
                Sq = scale[n, bdir, yq, xq]
                orig = messages[n, bdir, yq, xq, :] + np.log(Sq)[:, None]
                p = np.exp(orig) / Sq[:, None]
                G = in_grad[n, bdir, yq, xq, :]
                hor = (G + sgn * p * G.sum(axis=1, keepdims=True)) + \
                      (acc[:, 0] + sgn * p * acc[:, 0].sum(axis=1, keepdims=True))
                if cross:
                    up = saved[UP, yq, xq, :] + acc[:, 1]
                    down = saved[DOWN, yq, xq, :] + acc[:, 2]
                else:
                    up = np.zeros_like(hor)
                    down = np.zeros_like(hor)
                a = argmin[n, bdir, yq, xq, :]
                tot = hor + up + down
                P = len(ys)
                rows = np.repeat(np.arange(P), C)
                gu = np.zeros((P, C))
                np.add.at(gu, (rows, a.ravel()), tot.ravel())
                g_unary[n, ys, xs, :] += gu
                fv = mag(pw.f(n, gxy, ys, xs, cs[None, :], a))
                pw.add_grad(g_pw, n, gxy, ys, xs, cs[None, :], a, w[:, None] * tot)
                g_edge[n, gxy, ys, xs] += (fv * tot).sum(axis=1)
                tmp = np.zeros((P, 3, C))
                for j, comp in enumerate((hor, up, down)):
                    t = np.zeros((P, C))
                    np.add.at(t, (rows, a.ravel()), comp.ravel())
                    tmp[:, j] = t
                acc = tmp
                saved[bdir, ys, xs, :] = acc[:, 0]
    if pw.kind == "semantic_pixel":
        g_edge = np.zeros_like(g_edge)  # not computed (semantic_pw_pixel kernel :399-401)
    return g_unary, g_pw, g_edge, in_grad.copy()


# --------------------------------------------------------------------------------------
# message-passing layer around the op
# --------------------------------------------------------------------------------------
def softmin(x, T=1.0, axis=-1):
    z = -np.asarray(x, dtype=np.float64) / T
    z = z - z.max(axis=axis, keepdims=True)
    e = np.exp(z)
    return e / e.sum(axis=axis, keepdims=True)


def beliefs(cost, messages, T=1.0):
    """softmin((sum_s M_s + cost) / T) over labels (message_passing_op_cuda.py:170-173)."""
    return softmin(np.asarray(messages).sum(axis=1) + cost, T)
