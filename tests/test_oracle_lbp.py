"""Pins the FP64 LBP oracle (oracle/lbp.py) with checks independent of its own formulas."""
import itertools

import numpy as np
import pytest

from oracle import lbp as O
from _lbp_cases import (pack_stereo_jump, unpack_stereo_grad, random_stereo_problem, semantic_matrix,
                        pack_semantic_global, LEFT, RIGHT, UP, DOWN)


# ----------------------------------------------------------------------------- tie rule
def bitrev(i, bits):
    r = 0
    for _ in range(bits):
        r = (r << 1) | (i & 1)
        i >>= 1
    return r


def max_bitrev_rule(h):
    """Closed form of the tree order: among the minima, the label whose index, bit-reversed
    over log2(nextpow2(C)) bits, is largest."""
    C = len(h)
    bits = max(C - 1, 0).bit_length()
    m = h.min()
    cands = [l for l in range(C) if h[l] == m]
    return max(cands, key=lambda l: bitrev(l, bits))


@pytest.mark.parametrize("C", [1, 2, 3, 5, 8, 19, 25, 33, 64, 100, 192, 256])
def test_tree_argmin_equals_max_bitrev(C):
    rng = np.random.default_rng(C)
    for trial in range(200):
        h = rng.integers(0, 3, size=C).astype(np.float64)  # many ties
        v, l = O.tree_argmin(h[None, :])
        assert v[0] == h.min()
        assert l[0] == max_bitrev_rule(h), (h, l)


# ----------------------------------------------------------------------------- chains
def brute_min_marginals(unary, pair):
    """unary [L, C]; pair(x, l_x, l_{x+1}) energy of the edge (x, x+1)."""
    L, C = unary.shape
    mm = np.full((L, C), np.inf)
    for labels in itertools.product(range(C), repeat=L):
        e = sum(unary[i, labels[i]] for i in range(L))
        e += sum(pair(i, labels[i], labels[i + 1]) for i in range(L - 1))
        for i in range(L):
            mm[i, labels[i]] = min(mm[i, labels[i]], e)
    return mm


def _norm(b):
    return b - b.min(axis=-1, keepdims=True)


@pytest.mark.parametrize("C,delta", [(3, 3), (4, 10), (9, 3), (8, 2)])
def test_horizontal_chain_is_exact(C, delta):
    """On a chain min-sum BP is exact: cost + M_L + M_R are min-marginals up to a constant.
    Edge (x, x+1) parameters are shared by the RIGHT message from x and the LEFT message
    from x+1, which the packing makes consistent (RIGHT = transpose of LEFT)."""
    rng = np.random.default_rng(10 + C)
    N, H, W = 1, 2, 4 if C >= 8 else 5
    cost = rng.normal(size=(N, H, W, C))
    a = rng.uniform(0.05, 0.5, size=(W - 1, 2))
    L2 = np.maximum(rng.uniform(0.2, 1.0, size=W - 1), a.max(axis=1))
    e = rng.uniform(0.5, 1.5, size=(H, W - 1))
    aff = np.zeros((N, 4, 3, H, W))
    edge = np.ones((N, 4, H, W))
    for x in range(W - 1):
        aff[0, RIGHT, :, :, x] = np.array([a[x, 0], a[x, 1], L2[x]])[:, None]
        aff[0, LEFT, :, :, x + 1] = np.array([a[x, 0], a[x, 1], L2[x]])[:, None]
        edge[0, RIGHT, :, x] = e[:, x]
        edge[0, LEFT, :, x + 1] = e[:, x]
    jump = pack_stereo_jump(aff)
    M0 = np.zeros((N, 4, H, W, C))
    M, *_ = O.lbp_forward(cost, O.StereoPairwise(jump), edge, M0, delta)
    b = cost + M[:, LEFT] + M[:, RIGHT]

    for y in range(H):
        def pair(x, lx, lx1):
            # f_RIGHT(c=l_{x+1}, l=l_x): offset c-l = lx1-lx; LEFT/UP packing [o,L2,a0,0,a1,L2]
            dd = lx1 - lx
            J = jump[0, RIGHT, :, y, x]
            return J[np.clip(3 + dd, 1, 5)] * e[y, x]
        mm = brute_min_marginals(cost[0, y], pair)
        np.testing.assert_allclose(_norm(b[0, y]), _norm(mm), atol=1e-12)


def test_vertical_chain_is_exact_with_horizontal_unaries():
    """With W=1 the horizontal passes send nothing; the vertical passes then solve the
    column chain exactly."""
    rng = np.random.default_rng(3)
    N, H, W, C = 1, 5, 1, 4
    cost = rng.normal(size=(N, H, W, C))
    Mh = semantic_matrix(rng, C, zero_diag=False)
    Mv = semantic_matrix(rng, C, zero_diag=False)
    jump = pack_semantic_global(np.stack([Mh, Mv])[None])
    edge = rng.uniform(0.5, 1.5, size=(N, 4, H, W))
    # consistent edges: DOWN from y and UP from y+1 share a weight
    for y in range(H - 1):
        edge[0, UP, y + 1] = edge[0, DOWN, y]
    M0 = np.zeros((N, 4, H, W, C))
    M, *_ = O.lbp_forward(cost, O.SemanticGlobalPairwise(jump), edge, M0, delta=C)
    b = cost + M.sum(axis=1)

    def pair(y, ly, ly1):
        # DOWN from y: f(c=l_{y+1}, l=l_y) = jump[DOWN][l][c] = Mv^T[l_y][l_{y+1}] = Mv[l_{y+1}][l_y]
        return Mv[ly1, ly] * edge[0, DOWN, y, 0]
    mm = brute_min_marginals(cost[0, :, 0], pair)
    np.testing.assert_allclose(_norm(b[0, :, 0]), _norm(mm), atol=1e-12)


def test_semantic_pixel_chain_is_exact():
    rng = np.random.default_rng(4)
    N, H, W, C = 1, 1, 5, 3
    cost = rng.normal(size=(N, H, W, C))
    jump = np.zeros((4, H, W, C, C))
    for x in range(W - 1):
        A = rng.uniform(0, 1, size=(C, C))
        jump[RIGHT, 0, x] = A            # f_R(c,l) = A[l][c]
        jump[LEFT, 0, x + 1] = A.T       # f_L(c,l) = A^T[l][c] = A[c][l] = f_R(l,c)
    edge = np.ones((N, 4, H, W))
    M, *_ = O.lbp_forward(cost, O.SemanticPixelPairwise(jump), edge, np.zeros((N, 4, H, W, C)), delta=C)
    b = cost + M.sum(axis=1)

    def pair(x, lx, lx1):
        return jump[RIGHT, 0, x][lx, lx1]
    mm = brute_min_marginals(cost[0, 0], pair)
    np.testing.assert_allclose(_norm(b[0, 0]), _norm(mm), atol=1e-12)


# ----------------------------------------------------------------------------- invariants
def test_truncated_equals_full_on_ordered_pairwise():
    rng = np.random.default_rng(5)
    cost, aff, jump, edge = random_stereo_problem(rng, 1, 3, 6, 11, ordered=True)
    M0 = rng.normal(size=(1, 4, 3, 6, 11)) * 0.3
    pw = O.StereoPairwise(jump)
    Mt, at, *_ = O.lbp_forward(cost, pw, edge, M0, delta=3)
    Mf, af, *_ = O.lbp_forward(cost, pw, edge, M0, delta=11)
    np.testing.assert_allclose(Mt, Mf, atol=1e-12)


def test_logsumexp_of_written_messages_is_zero():
    rng = np.random.default_rng(6)
    cost, aff, jump, edge = random_stereo_problem(rng, 2, 4, 5, 6, ordered=False)
    M, a, S, fg, ls = O.lbp_forward(cost, O.StereoPairwise(jump), edge, np.zeros((2, 4, 4, 5, 6)), 3)
    lse = np.log(np.exp(M).sum(-1))
    written = np.ones((2, 4, 4, 5), dtype=bool)
    written[:, LEFT, :, 0] = False
    written[:, RIGHT, :, -1] = False
    written[:, UP, 0, :] = False
    written[:, DOWN, -1, :] = False
    np.testing.assert_allclose(lse[written], 0.0, atol=1e-12)
    assert np.all(M[~written] == 0.0)


def test_message_values_are_attained_by_argmin():
    """msg(c) = h(a(c)) + f(c, a(c)) w (or the truncation term); checked by recomputing h."""
    rng = np.random.default_rng(7)
    C = 9
    cost, aff, jump, edge = random_stereo_problem(rng, 1, 3, 4, C, ordered=False)
    pw = O.StereoPairwise(jump)
    h = cost[0, :, 0, :].copy()
    ys, xs = np.arange(3), np.zeros(3, dtype=int)
    w = edge[0, RIGHT, :, 0]
    mn, a, S, fg, ls = O.send(h, pw, 0, RIGHT, ys, xs, w, 3)
    msg = mn + np.log(S)[:, None]
    cs = np.arange(C)
    fv = np.where(fg, pw.trunc(0, RIGHT, ys, xs)[:, None], pw.f(0, RIGHT, ys, xs, cs[None, :], a))
    np.testing.assert_allclose(msg, h[np.arange(3)[:, None], a] + fv * w[:, None], atol=1e-12)
    # and it is the minimum over all candidates of the truncated rule
    J = jump[0, RIGHT][:, ys, xs].T
    for p in range(3):
        for c in range(C):
            cands = [h[p].min() + J[p, 5] * w[p]]
            cands += [h[p, l] + J[p, np.clip(3 + c - l, 1, 5)] * w[p] for l in range(max(c - 3, 0), min(c + 4, C))]
            assert abs(min(cands) - msg[p, c]) < 1e-12


@pytest.mark.parametrize("axis", ["x", "y"])
def test_mirror_symmetry(axis):
    """Flipping the image along x swaps the LEFT/RIGHT passes (slots, edges and pairwise
    terms); along y it swaps UP/DOWN.  Label order is unchanged, so argmins must agree."""
    rng = np.random.default_rng(8)
    N, H, W, C = 1, 4, 5, 7
    cost, aff, jump, edge = random_stereo_problem(rng, N, H, W, C, ordered=False)
    M0 = rng.normal(size=(N, 4, H, W, C)) * 0.2
    M, a, *_ = O.lbp_forward(cost, O.StereoPairwise(jump), edge, M0, 3)
    if axis == "x":
        fl = lambda t, ax: np.flip(t, axis=ax)
        sw = [RIGHT, LEFT, UP, DOWN]
        cost_f = fl(cost, 2)
        M0_f = fl(M0, 3)[:, sw]
        jump_f = fl(jump, 4)[:, sw]
        edge_f = fl(edge, 3)[:, sw]
        Mf, af, *_ = O.lbp_forward(cost_f, O.StereoPairwise(jump_f), edge_f, M0_f, 3)
        np.testing.assert_allclose(fl(Mf, 3)[:, sw], M, atol=1e-12)
        np.testing.assert_array_equal(fl(af, 3)[:, sw], a)
    else:
        fl = lambda t, ax: np.flip(t, axis=ax)
        sw = [LEFT, RIGHT, DOWN, UP]
        cost_f = fl(cost, 1)
        M0_f = fl(M0, 2)[:, sw]
        jump_f = fl(jump, 3)[:, sw]
        edge_f = fl(edge, 2)[:, sw]
        Mf, af, *_ = O.lbp_forward(cost_f, O.StereoPairwise(jump_f), edge_f, M0_f, 3)
        np.testing.assert_allclose(fl(Mf, 2)[:, sw], M, atol=1e-12)
        np.testing.assert_array_equal(fl(af, 2)[:, sw], a)


# ----------------------------------------------------------------------------- gradients
def _loss_and_grad(rng, shape_msgs, shape_cost):
    Wm = rng.normal(size=shape_msgs)
    Wb = rng.normal(size=shape_cost)

    def loss(cost, M):
        return float((Wm * M).sum() + (Wb * O.beliefs(cost, M)).sum())

    def grads(cost, M):
        # d/dM and d/dcost of the loss, analytically (softmin Jacobian)
        b = O.beliefs(cost, M)
        gb = Wb
        gz = b * (gb - (gb * b).sum(-1, keepdims=True))   # softmax grad wrt z = -(sum M + cost)
        gE = -gz
        gM = Wm + gE[:, None]
        return gE, gM
    return loss, grads


def _fd_check(fun, x, analytic, rng, n_probe=40, eps=1e-6, rtol=1e-5, atol=1e-7):
    flat = x.reshape(-1)
    idx = rng.choice(flat.size, size=min(n_probe, flat.size), replace=False)
    for i in idx:
        old = flat[i]
        flat[i] = old + eps
        lp = fun()
        flat[i] = old - eps
        lm = fun()
        flat[i] = old
        fd = (lp - lm) / (2 * eps)
        an = analytic.reshape(-1)[i]
        assert abs(fd - an) <= atol + rtol * max(abs(fd), abs(an)), (i, fd, an)


@pytest.mark.parametrize("variant,iters,ordered", [("stereo", 1, True), ("stereo", 2, False),
                                                   ("stereo5", 2, True),
                                                   ("stereo_full", 2, True),
                                                   ("semantic_global", 2, True),
                                                   ("semantic_pixel", 2, True)])
def test_exact_adjoint_matches_finite_differences(variant, iters, ordered):
    """FP64 central differences of a loss on the final messages and the beliefs.  Stereo
    pairwise terms are perturbed through the affinity packing: the packed truncation entry
    and the far entries share one value, and moving them together is what is
    differentiable (moving one alone breaks an exact tie between window and global terms)."""
    rng = np.random.default_rng(20 + iters)
    N, H, W, C = 1, 3, 4, 5
    aff = None
    if variant.startswith("stereo"):
        C = 7 if variant != "stereo_full" else 5
        K_in = 5 if variant == "stereo5" else 3
        cost, aff, jump, edge = random_stereo_problem(rng, N, H, W, C, K_in=K_in, ordered=ordered)
        kind = "stereo"
        delta = C if variant == "stereo_full" else 3
    elif variant == "semantic_global":
        cost = rng.normal(size=(N, H, W, C))
        jump = pack_semantic_global(np.stack([semantic_matrix(rng, C), semantic_matrix(rng, C)])[None])
        edge = rng.uniform(0.5, 1.5, size=(N, 4, H, W))
        kind, delta = "semantic_global", C
    else:
        cost = rng.normal(size=(N, H, W, C))
        jump = rng.uniform(0, 1, size=(4, H, W, C, C))
        edge = rng.uniform(0.5, 1.5, size=(N, 4, H, W))
        kind, delta = "semantic_pixel", C
    M0 = rng.normal(size=(N, 4, H, W, C)) * 0.3
    loss, grads = _loss_and_grad(rng, M0.shape, cost.shape)

    def cur_jump():
        return pack_stereo_jump(aff) if aff is not None else jump

    def run():
        return O.lbp_forward_iters(cost, O.make_pairwise(kind, cur_jump()), edge, M0, delta, iters)

    def fwd_loss():
        M, _ = run()
        return loss(cost, M)

    M, recs = run()
    gE, gM = grads(cost, M)
    gc, gp, ge, gm0 = O.lbp_backward_exact_iters(cost, O.make_pairwise(kind, cur_jump()), edge, recs, gM)
    gc = gc + gE
    _fd_check(fwd_loss, cost, gc, rng)
    _fd_check(fwd_loss, edge, ge, rng)
    _fd_check(fwd_loss, M0, gm0, rng)
    if aff is not None:
        _fd_check(fwd_loss, aff, unpack_stereo_grad(gp, aff.shape[2]), rng, n_probe=60)
    else:
        _fd_check(fwd_loss, jump, gp, rng, n_probe=60)


@pytest.mark.parametrize("variant", ["stereo", "stereo_full", "semantic_global", "semantic_pixel"])
def test_reference_backward_equals_exact_at_one_iteration(variant):
    """Mode R (the reference's backward) equals the exact adjoint for cost, edge and
    pairwise gradients after one iteration: its un-normalised cross terms sum to zero.
    The stereo truncation bins (1 and K-1 hold the same value after packing) are folded."""
    rng = np.random.default_rng(31)
    N, H, W, C = 2, 3, 4, 6
    if variant.startswith("stereo"):
        cost, aff, jump, edge = random_stereo_problem(rng, N, H, W, C, ordered=True)
        kind, delta = "stereo", (3 if variant == "stereo" else C)
    elif variant == "semantic_global":
        cost = rng.normal(size=(N, H, W, C))
        jump = pack_semantic_global(np.stack([semantic_matrix(rng, C), semantic_matrix(rng, C)])[None])
        edge = rng.uniform(0.5, 1.5, size=(N, 4, H, W))
        kind, delta = "semantic_global", C
    else:
        cost = rng.normal(size=(N, H, W, C))
        jump = rng.uniform(0, 1, size=(4, H, W, C, C))
        edge = rng.uniform(0.5, 1.5, size=(N, 4, H, W))
        kind, delta = "semantic_pixel", C
    pw = O.make_pairwise(kind, jump)
    M0 = np.zeros((N, 4, H, W, C))
    M, a, S, fg, ls = O.lbp_forward(cost, pw, edge, M0, delta)
    G = rng.normal(size=M.shape)
    rc, rp, re, rm = O.lbp_backward_reference(cost, pw, edge, G, M, a, S)
    ec, ep, ee, em = O.lbp_backward_exact(cost, pw, edge, M, a, fg, G)
    np.testing.assert_allclose(rc, ec, atol=1e-10)
    if kind != "semantic_pixel":
        np.testing.assert_allclose(re, ee, atol=1e-10)
    if kind == "stereo":
        rp = rp.copy(); ep = ep.copy()
        rp[:, :, -1] += rp[:, :, 1]; rp[:, :, 1] = 0
        ep[:, :, -1] += ep[:, :, 1]; ep[:, :, 1] = 0
    np.testing.assert_allclose(rp, ep, atol=1e-10)
    np.testing.assert_array_equal(rm, G)


# --------------------------------------------------------------------------------------
# the parity bars' own tools: the FP64 step replay with its fp32 tie mask, and the
# condition magnitudes of the adjoint
# --------------------------------------------------------------------------------------
def test_local_argmins_reproduce_the_oracle_and_an_independent_fp32_implementation():
    """On the oracle's own messages the step replay gives the oracle's argmins; on the C++
    fp32 restatement's messages it gives that implementation's argmins wherever it calls the
    decision clear (an fp32 evaluation with its own summation order)."""
    from oracle import cpu
    from _lbp_cases import random_stereo_problem
    rng = np.random.default_rng(41)
    for C, delta in ((24, 3), (9, 9)):
        cost, aff, jump, edge = random_stereo_problem(rng, 1, 4, 30, C)
        f32 = lambda a: np.asarray(a, dtype=np.float32)
        cost, jump, edge = f32(cost), f32(jump), f32(edge)
        M0 = f32(rng.normal(size=(1, 4, 4, 30, C)) * 0.3)
        pw = O.StereoPairwise(jump)
        Mo, ao, so, fo, lso = O.lbp_forward(cost, pw, edge, M0, delta)
        lab, fg, clear, wr = O.local_argmins(cost, pw, edge, M0, f32(Mo), delta)
        m = clear & wr[..., None]
        assert m.sum() > 0.99 * wr.sum() * C
        assert np.array_equal(lab[m], np.where(fo, lso[..., None], ao)[m]) and np.array_equal(fg[m], fo[m])
        M32, a32, f32_ = cpu.lbp_forward(cost, jump, edge, M0, delta, dtype=np.float32)
        lab, fg, clear, wr = O.local_argmins(cost, pw, edge, M0, M32, delta)
        m = clear & wr[..., None]
        assert np.array_equal(lab[m], a32[m]) and np.array_equal(fg[m], f32_[m].astype(bool))


def test_local_argmins_flag_rounding_ties_but_not_identical_operands():
    """Two window candidates one ulp apart are a rounding tie (not clear); candidates from
    bitwise-identical operands are an exact tie every implementation resolves by the
    reference's '<=' order (clear, last offset wins)."""
    C, W = 8, 2
    jump = np.zeros((1, 4, 6, 1, W), dtype=np.float32)
    jump[:, :, 1] = jump[:, :, 5] = 5.0          # truncation far away
    jump[:, :, 2] = jump[:, :, 4] = 0.25         # P1
    edge = np.ones((1, 4, 1, W), dtype=np.float32)
    M0 = np.zeros((1, 4, 1, W, C), dtype=np.float32)
    pw = O.StereoPairwise(jump)
    for kind in ("ulp", "identical"):
        cost = np.full((1, 1, W, C), 1.0, dtype=np.float32)
        cost[0, 0, 0, 3] = 0.0                    # sender x=0: minimum at label 3
        if kind == "ulp":
            cost[0, 0, 0, 5] = np.nextafter(np.float32(0.0), np.float32(1.0)) * 0 + np.float32(1e-7)
            cost[0, 0, 0, 3] = 0.0
        else:
            cost[0, 0, 0, 5] = 0.0                # labels 3 and 5 identical
        Mo, ao, *_ = O.lbp_forward(cost, pw, edge, M0, 3)
        lab, fg, clear, wr = O.local_argmins(cost, pw, edge, M0, Mo.astype(np.float32), 3)
        # receiver label 4 sees sender labels 3 and 5 through the same P1 entry
        r = clear[0, 0, 0, 1, 4]                   # slot LEFT at x=1 (RIGHT pass)
        if kind == "ulp":
            assert not r
        else:
            assert r and lab[0, 0, 0, 1, 4] == 5 and ao[0, 0, 0, 1, 4] == 5


def test_condition_magnitudes_bound_the_gradients_and_an_fp32_evaluation():
    """kappa (absolute=True) dominates |gradient| elementwise, is linear in |G|, and the C++
    fp32 restatement's exact backward errs by well under 1e-5 kappa on a realistic problem."""
    from oracle import cpu
    from _lbp_cases import random_stereo_problem
    rng = np.random.default_rng(43)
    cost, aff, jump, edge = random_stereo_problem(rng, 1, 5, 40, 32)
    f32 = lambda a: np.asarray(a, dtype=np.float32)
    cost, jump, edge = f32(cost), f32(jump), f32(edge)
    M0 = np.zeros((1, 4, 5, 40, 32), dtype=np.float32)
    pw = O.StereoPairwise(jump)
    Mo, ao, so, fo, lso = O.lbp_forward(cost, pw, edge, M0, 3)
    G = rng.normal(size=Mo.shape)
    g = O.lbp_backward_exact(cost, pw, edge, Mo, ao, fo, G)
    k = O.lbp_backward_exact(cost, pw, edge, Mo, ao, fo, G, absolute=True)
    k2 = O.lbp_backward_exact(cost, pw, edge, Mo, ao, fo, -2 * G, absolute=True)
    for a, b, c in zip(g, k, k2):
        assert np.all(np.abs(a) <= b * (1 + 1e-12) + 1e-300)
        np.testing.assert_allclose(c, 2 * b, rtol=1e-12)
    lab = np.where(fo, lso[..., None], ao).astype(np.int32)
    r32 = cpu.lbp_backward_exact(jump, edge, f32(G), f32(Mo), lab, fo.astype(np.uint8), 3, dtype=np.float32)
    for a32, a64, kk in zip(r32, g, k):
        nz = kk > 0
        assert (np.abs(a32.astype(np.float64) - a64)[nz] / kk[nz]).max() < 1e-5
