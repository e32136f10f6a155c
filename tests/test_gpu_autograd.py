"""GPU tests of the autograd layer (bplayers_amd.ops.*) against the FP64 oracle."""
import numpy as np
import pytest
import torch

import _parity as P
from oracle import lbp as O
from _lbp_cases import pack_stereo_jump, unpack_stereo_grad, semantic_matrix, pack_semantic_global

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _t(a, grad=False):
    return torch.tensor(np.asarray(a, dtype=np.float32), device=DEV, requires_grad=grad)


def oracle_pipeline(prob, aff, edge, iters, delta, Wb, T=1.0):
    """loss = sum(Wb * softmin(sum messages + cost)) with cost = -prob, messages from zeros."""
    cost = -np.asarray(prob, dtype=np.float64)
    jump = pack_stereo_jump(aff)
    M = np.zeros(cost.shape[:1] + (4,) + cost.shape[1:])
    M, recs = O.lbp_forward_iters(cost, O.StereoPairwise(jump), edge, M, delta, iters)
    return float((Wb * O.beliefs(cost, M, T)).sum()), M, recs


def oracle_pipeline_grads(prob, aff, edge, iters, delta, Wb, absolute=False):
    """Analytic FP64 gradients of oracle_pipeline's loss w.r.t. (prob, affinities, edge)
    (softmin Jacobian, the exact adjoint chained over the iterations, the packing adjoint);
    ``absolute``: their condition magnitudes."""
    cost = -np.asarray(prob, dtype=np.float64)
    _, M, recs = oracle_pipeline(prob, aff, edge, iters, delta, Wb)
    gE, kE = P.loss_grad_of_beliefs(cost + M.sum(axis=1), Wb)
    g = kE if absolute else gE
    gM = np.broadcast_to(g[:, None], M.shape)
    gc, gj, ge, _ = O.lbp_backward_exact_iters(cost, O.StereoPairwise(pack_stereo_jump(aff)), edge, recs, gM,
                                               absolute=absolute)
    gprob = (gc + g) if absolute else -(gc + g)
    ga = unpack_stereo_grad(gj, aff.shape[2])
    return gprob, np.abs(ga) if absolute else ga, ge


@pytest.mark.parametrize("iters", [1, 3])
def test_stereo_message_passing_exact_gradients_match_fd(iters):
    from bplayers_amd.ops.lbp_stereo.message_passing_op_cuda import MessagePassing
    rng = np.random.default_rng(iters)
    N, H, W, C = 1, 4, 6, 12
    prob = rng.dirichlet(np.ones(C), size=(N, H, W)).astype(np.float32)
    p1 = rng.uniform(0.05, 0.3, size=(N, 4, H, W))
    p2 = rng.uniform(0.4, 0.8, size=(N, 4, H, W))
    aff = np.stack([p1, p1, p2], axis=2).astype(np.float32)
    edge = rng.uniform(0.5, 1.5, size=(N, 4, H, W)).astype(np.float32)
    Wb = rng.normal(size=(N, H, W, C))
    mp = MessagePassing(DEV, iters, C, 3, grad_mode="exact")
    tp, ta, te = _t(prob, True), _t(aff, True), _t(edge, True)
    m0 = torch.zeros((N, 4, H, W, C), device=DEV)
    beliefs, msgs = mp(tp, te, ta, None, m0)
    loss = (beliefs * torch.tensor(Wb, dtype=torch.float32, device=DEV)).sum()
    loss.backward()
    L0, Mo, _ = oracle_pipeline(prob, aff, edge, iters, 3, Wb)
    P.elementwise("loss", np.array(float(loss.detach())), np.array(L0))
    ora = oracle_pipeline_grads(prob, aff, edge, iters, 3, Wb)
    for name, g, o in zip(("grad prob", "grad affinities", "grad edge"), (tp.grad, ta.grad, te.grad), ora):
        P.normwise(name, g.cpu().numpy(), o)
    # the analytic oracle gradient against FP64 central differences of the oracle pipeline
    eps = 1e-5
    for name, arr, g in (("prob", prob, tp.grad), ("aff", aff, ta.grad), ("edge", edge, te.grad)):
        a64 = arr.astype(np.float64)
        flat = a64.reshape(-1)
        gg = g.cpu().numpy().reshape(-1)
        for i in rng.choice(flat.size, size=25, replace=False):
            o = flat[i]
            flat[i] = o + eps
            lp = oracle_pipeline(a64.reshape(arr.shape) if name == "prob" else prob,
                                 a64.reshape(arr.shape) if name == "aff" else aff,
                                 a64.reshape(arr.shape) if name == "edge" else edge, iters, 3, Wb)[0]
            flat[i] = o - eps
            lm = oracle_pipeline(a64.reshape(arr.shape) if name == "prob" else prob,
                                 a64.reshape(arr.shape) if name == "aff" else aff,
                                 a64.reshape(arr.shape) if name == "edge" else edge, iters, 3, Wb)[0]
            flat[i] = o
            fd = (lp - lm) / (2 * eps)
            og = {"prob": ora[0], "aff": ora[1], "edge": ora[2]}[name].reshape(-1)[i]
            assert abs(fd - og) <= 1e-6 * max(1.0, abs(fd)), (name, i, fd, og)
            assert abs(fd - gg[i]) <= 1e-4 * max(1.0, abs(fd)), (name, i, fd, gg[i])


def test_stereo_reference_mode_multi_iteration_matches_restatement():
    """grad_mode='reference', 2 iterations: every iteration's backward sees the final
    (aliased) messages and in_grad of the final messages (the reference's semantics)."""
    from bplayers_amd.ops.lbp_stereo.message_passing_op_cuda import LBPMinSumFunction
    rng = np.random.default_rng(11)
    N, H, W, C = 1, 5, 7, 16
    cost = rng.normal(size=(N, H, W, C)).astype(np.float32)
    p1 = rng.uniform(0.05, 0.3, size=(N, 4, H, W))
    p2 = rng.uniform(0.4, 0.8, size=(N, 4, H, W))
    aff = np.stack([p1, p1, p2], axis=2).astype(np.float32)
    edge = rng.uniform(0.5, 1.5, size=(N, 4, H, W)).astype(np.float32)
    G = rng.normal(size=(N, 4, H, W, C))
    tc, ta, te = _t(cost, True), _t(aff, True), _t(edge, True)
    m = torch.zeros((N, 4, H, W, C), device=DEV, requires_grad=True)
    off = torch.zeros((N, 4, H, W), device=DEV)
    x = m
    for _ in range(2):
        x = LBPMinSumFunction.apply(tc, te, x, 3, ta, off)
    (x * torch.tensor(G, dtype=torch.float32, device=DEV)).sum().backward()
    # restatement: per-iteration argmins/scales, final messages, in_grad = G for both
    jump = pack_stereo_jump(aff)
    pw = O.StereoPairwise(jump)
    M = np.zeros((N, 4, H, W, C))
    recs = []
    for _ in range(2):
        Mn, a, S, fg, ls = O.lbp_forward(cost, pw, edge, M, 3)
        recs.append((a, S))
        M = Mn
    gc = np.zeros_like(cost, dtype=np.float64)
    gp = np.zeros_like(jump)
    ge = np.zeros_like(edge, dtype=np.float64)
    for a, S in recs:
        c_, p_, e_, _ = O.lbp_backward_reference(cost, pw, edge, G, M, a, S)
        gc += c_
        gp += p_
        ge += e_
    for name, g, o in zip(("grad cost", "grad affinities", "grad edge"), (tc.grad, ta.grad, te.grad),
                          (gc, unpack_stereo_grad(gp, 3), ge)):
        P.normwise(name, g.cpu().numpy(), o)
    assert np.array_equal(m.grad.cpu().numpy(), G.astype(np.float32))


@pytest.mark.parametrize("shape", [(1, 6, 37, 64), (2, 5, 23, 100), (1, 5, 9, 12)])
def test_stereo_exact_chain_equals_per_iteration_functions(shape):
    """One chained autograd node over 3 iterations == 3 per-iteration exact Functions: same
    messages, and the same gradients up to the order in which the iterations' terms are summed
    (paired fast kernel for C > 16 with delta 3; generic kernels for C <= 16)."""
    from bplayers_amd.ops.lbp_stereo.message_passing_op_cuda import (LBPMinSumExactChainFunction,
                                                                      LBPMinSumExactFunction)
    N, H, W, C = shape
    rng = np.random.default_rng(C + W)
    cost = rng.normal(size=(N, H, W, C))
    aff = np.log1p(np.exp(rng.normal(size=(N, 4, 3, H, W))))
    edge = rng.uniform(0.5, 1.5, size=(N, 4, H, W))
    m0 = rng.normal(size=(N, 4, H, W, C)) * 0.3
    Wm = torch.tensor(rng.normal(size=(N, 4, H, W, C)), dtype=torch.float32, device=DEV)
    offset = torch.zeros((N, 4, H, W), device=DEV)
    outs, grads = [], []
    for chained in (True, False):
        tc, ta, te, tm = _t(cost, True), _t(aff, True), _t(edge, True), _t(m0, True)
        if chained:
            m = LBPMinSumExactChainFunction.apply(tc, te, tm, 3, ta, offset, 3)
        else:
            m = tm
            for _ in range(3):
                m = LBPMinSumExactFunction.apply(tc, te, m, 3, ta, offset)
        (m * Wm).sum().backward()
        torch.cuda.synchronize()
        outs.append(m.detach())
        grads.append([x.grad for x in (tc, ta, te, tm)])
    assert torch.equal(outs[0], outs[1])
    for name, a, b in zip(("cost", "affinities", "edge", "messages"), *grads):
        torch.testing.assert_close(a, b, rtol=1e-5, atol=1e-5, msg=lambda m_: f"grad {name}: {m_}")


@pytest.mark.parametrize("variant", ["global", "pixel"])
def test_semantic_message_passing_exact_gradients(variant):
    rng = np.random.default_rng(5)
    N, H, W, C = 2, 4, 5, 7
    prob = rng.dirichlet(np.ones(C), size=(N, H, W)).astype(np.float32)
    edge = rng.uniform(0.5, 1.5, size=(N, 4, H, W)).astype(np.float32)
    Wb = rng.normal(size=(N, H, W, C))
    if variant == "global":
        from bplayers_amd.ops.lbp_semantic_pw.message_passing_op_cuda import MessagePassing
        jump = np.stack([semantic_matrix(rng, C), semantic_matrix(rng, C)])[None].astype(np.float32)
        packed = pack_semantic_global(jump)
        pwk = "semantic_global"
    else:
        from bplayers_amd.ops.lbp_semantic_pw_pixel.message_passing_op_pw_pixel import MessagePassing
        jump = rng.uniform(0, 1, size=(2, H, W, C, C)).astype(np.float32)
        packed = np.zeros((4, H, W, C, C))
        packed[0] = jump[0]
        packed[1, :, 1:] = np.swapaxes(jump[0, :, :-1], -1, -2)
        packed[2] = jump[1]
        packed[3, 1:] = np.swapaxes(jump[1, :-1], -1, -2)
        pwk = "semantic_pixel"
    mp = MessagePassing(DEV, 2, C, C, grad_mode="exact")
    tp, tj, te = _t(prob, True), _t(jump, True), _t(edge, True)
    m0 = torch.zeros((N, 4, H, W, C), device=DEV)
    beliefs = mp(tp, te, m0, tj) if variant == "pixel" else mp(tp, te, m0, tj, False)
    (beliefs * torch.tensor(Wb, dtype=torch.float32, device=DEV)).sum().backward()
    cost = -prob.astype(np.float64)
    M, recs = O.lbp_forward_iters(cost, O.make_pairwise(pwk, packed), edge, np.zeros((N, 4, H, W, C)), C, 2)
    gE, _ = P.loss_grad_of_beliefs(cost + M.sum(axis=1), Wb)
    gM = np.broadcast_to(gE[:, None], M.shape)
    gc, gpk, ge, _ = O.lbp_backward_exact_iters(cost, O.make_pairwise(pwk, packed), edge, recs, gM)
    if variant == "global":
        gj = np.stack([gpk[0, 0] + gpk[0, 1].T, gpk[0, 2] + gpk[0, 3].T])[None]
    else:
        gj = np.zeros_like(jump, dtype=np.float64)
        gj[0] = gpk[0]
        gj[0, :, :-1] += np.swapaxes(gpk[1, :, 1:], -1, -2)
        gj[1] = gpk[2]
        gj[1, :-1] += np.swapaxes(gpk[3, 1:], -1, -2)
    for name, g, o in zip(("grad prob", "grad edge", "grad pairwise"), (tp.grad, te.grad, tj.grad),
                          (-(gc + gE), ge, gj)):
        P.normwise(name, g.cpu().numpy(), o)


def test_flow_bp_pipeline_runs_and_matches_oracle_forward():
    """FlowMpSadFunction -> softmax -> BP (5-D path: u then v from u's messages)."""
    from bplayers_amd.ops.flow_mp_sad.flow_mp_sad import FlowMpSadFunction
    from bplayers_amd.ops.lbp_stereo.bp_op_cuda import BP
    from oracle import sad as S
    rng = np.random.default_rng(3)
    f0 = rng.normal(size=(1, 8, 12, 20)).astype(np.float32)
    f1 = rng.normal(size=(1, 8, 12, 20)).astype(np.float32)
    t0, t1 = _t(f0, True), _t(f1, True)
    cu, cv, us, vs = FlowMpSadFunction.apply(t0, t1, 8)
    pu = torch.softmax(-cu, dim=3)
    pv = torch.softmax(-cv, dim=3)
    prob = torch.cat((pu.unsqueeze(1), pv.unsqueeze(1)), dim=1).contiguous()
    bp = BP(DEV, max_iter=1, num_labels=9, delta=3, mode_inference="wta")
    aff = torch.zeros((1, 4, 3, 12, 20), device=DEV)
    aff[:, :, :2] = 0.1
    aff[:, :, 2] = 0.5
    flow, beliefs, messages = bp(prob, None, aff, None)
    beliefs.sum().backward()
    assert t0.grad is not None and torch.isfinite(t0.grad).all()
    # oracle forward of the same chain
    ocu, ocv, _, _ = S.flow_sad_forward(f0, f1, 8)
    sm = lambda x: np.exp(-x - (-x).max(-1, keepdims=True)) / np.exp(-x - (-x).max(-1, keepdims=True)).sum(-1, keepdims=True)
    jump = pack_stereo_jump(aff.cpu().numpy())
    edge = np.ones((1, 4, 12, 20))
    Mu, *_ = O.lbp_forward(-sm(ocu), O.StereoPairwise(jump), edge, np.zeros((1, 4, 12, 20, 9)), 3)
    Mv, *_ = O.lbp_forward(-sm(ocv), O.StereoPairwise(jump), edge, Mu, 3)
    P.elementwise("beliefs v", beliefs[:, 1].detach().cpu().numpy(), O.beliefs(-sm(ocv), Mv))
    P.elementwise("beliefs u", beliefs[:, 0].detach().cpu().numpy(), O.beliefs(-sm(ocu), Mu))


def test_flow_matching_layer_matches_oracle():
    """FlowMatchingSad: softmax(-min-projected SAD / T) of both flow components, stacked as
    [N,2,H,W,K], and the feature gradients of a weighted sum of it, against the FP64 oracle
    (flow SAD forward/backward restatement plus the softmax Jacobian)."""
    from bplayers_amd.flow_matching import FlowMatchingSad
    from oracle import sad as S
    rng = np.random.default_rng(9)
    N, Cf, H, W, sws, T = 1, 6, 10, 14, 6, 1.7
    f0 = rng.normal(size=(N, Cf, H, W)).astype(np.float32)
    f1 = rng.normal(size=(N, Cf, H, W)).astype(np.float32)
    K = sws + 1
    Wt = rng.normal(size=(N, 2, H, W, K))
    m = FlowMatchingSad(sws, init_temp=T, device=DEV)
    t0, t1 = _t(f0, True), _t(f1, True)
    prob = m(t0, t1)
    (prob * torch.tensor(Wt, dtype=torch.float32, device=DEV)).sum().backward()
    cu, cv, us, vs = S.flow_sad_forward(f0.astype(np.float64), f1.astype(np.float64), sws)

    def softmax(x):
        e = np.exp(x - x.max(-1, keepdims=True))
        return e / e.sum(-1, keepdims=True)
    pu, pv = softmax(-cu / T), softmax(-cv / T)
    P.elementwise("probabilities", prob.detach().cpu().numpy(), np.stack([pu, pv], axis=1))
    # d/dcv of sum(W * softmax(-cv/T)) = -(1/T) p (W - sum(W p)), and its condition magnitudes
    gu, ku = P.loss_grad_of_beliefs(cu, Wt[:, 0], T)
    gv, kv = P.loss_grad_of_beliefs(cv, Wt[:, 1], T)
    d0, d1 = S.flow_sad_backward(f0.astype(np.float64), f1.astype(np.float64), sws, gu, gv, us, vs)
    k0, k1 = S.flow_sad_backward(f0.astype(np.float64), f1.astype(np.float64), sws, ku, kv, us, vs, absolute=True)
    P.conditioned("df0", t0.grad.cpu().numpy(), d0, k0)
    P.conditioned("df1", t1.grad.cpu().numpy(), d1, k1)


def test_multiscale_bp_matches_oracle():
    """Coarse-to-fine BP over a 3-level pyramid (stereo.multiscale_bp) against the same loop on
    the FP64 oracle: trilinear x2 upsampling of the coarser beliefs, half of them added to the
    finer probabilities (in place, without gradients), BP with the default P1/P2 per level;
    the returned lists keep the reference's order [level 0, level 2, level 1]."""
    import torch.nn.functional as F
    from bplayers_amd.ops.lbp_stereo.bp_op_cuda import BP
    from bplayers_amd.stereo import multiscale_bp
    rng = np.random.default_rng(12)
    shapes = [(1, 16, 24, 16), (1, 8, 12, 8), (1, 4, 6, 4)]
    probs = [rng.dirichlet(np.ones(s[-1]), size=s[:3]) for s in shapes]
    pv = [torch.tensor(p, dtype=torch.float32, device=DEV) for p in probs]
    crfs = [[BP(DEV, max_iter=1, num_labels=s[-1], delta=3, mode_inference='wta')] for s in shapes]
    with torch.no_grad():
        disps, beliefs = multiscale_bp(pv, crfs, min_disps=[0, 0, 0])
    torch.cuda.synchronize()
    assert [tuple(b.shape) for b in beliefs] == [shapes[0], shapes[2], shapes[1]]
    b_in, ob, oe = None, {}, {}
    for lvl in (2, 1, 0):
        N, H, W, D = shapes[lvl]
        p = probs[lvl].astype(np.float64)
        if b_in is not None:
            up = F.interpolate(torch.from_numpy(b_in)[:, None], scale_factor=2.0, mode='trilinear')[:, 0].numpy()
            p = p + up / 2.0
            np.testing.assert_allclose(pv[lvl].cpu().numpy(), p, rtol=1e-5, atol=1e-6)  # the in-place addition
        aff = np.zeros((N, 4, 3, H, W))
        aff[:, :, :2] = 0.1
        aff[:, :, 2] = 2.5
        cost = -p
        M, _ = O.lbp_forward_iters(cost, O.StereoPairwise(pack_stereo_jump(aff)), np.ones((N, 4, H, W)),
                                   np.zeros((N, 4, H, W, D)), 3, 1)
        ob[lvl] = b_in = O.beliefs(cost, M)
        oe[lvl] = cost + M.sum(axis=1)
    for i, lvl in enumerate((0, 2, 1)):   # level l: 3 - l solves chained through the upsampled priors
        P.elementwise(f"beliefs level {lvl}", beliefs[i].cpu().numpy(), ob[lvl], tol=P.beliefs_bar(3 - lvl))
    # winner-take-all: identical where the energy minimum leads by more than twice the largest
    # deviation of the GPU's beliefs, read as energies (log b differs from -E by a constant)
    dev = P.energy_deviation(beliefs[0].cpu().numpy(), oe[0])
    clear = P.winner_clear(oe[0], dev)
    print(f"    WTA: energy deviation up to {dev.max():.3g}, {1 - clear.mean():.2e} of pixels within twice their "
          f"own deviation of a tie", flush=True)
    assert clear.mean() > 0.99
    d0 = disps[0].cpu().numpy()[:, 0]
    assert np.array_equal(d0[clear], ob[0].argmax(-1)[clear])


def test_multiscale_flow_bp_matches_oracle():
    """Coarse-to-fine flow BP over 2 levels (flow.multiscale_flow_bp) against the same loop on
    the FP64 oracle: per level, u from zero messages and v from u's messages (the 5-D BP path),
    beliefs of each component upsampled trilinearly to (2H, 2W, 2K-1) and half of them added to
    the finer probabilities; flow = winner-take-all label - sws // 2."""
    import torch.nn.functional as F
    from bplayers_amd.ops.lbp_stereo.bp_op_cuda import BP
    from bplayers_amd.flow import multiscale_flow_bp
    rng = np.random.default_rng(13)
    shapes = [(1, 2, 8, 12, 9), (1, 2, 4, 6, 5)]
    sws = [8, 4]
    probs = [rng.dirichlet(np.ones(s[-1]), size=s[:4]) for s in shapes]
    pv = [torch.tensor(p, dtype=torch.float32, device=DEV) for p in probs]
    crfs = [[BP(DEV, max_iter=1, num_labels=s[-1], delta=3, mode_inference='wta')] for s in shapes]
    with torch.no_grad():
        flows, beliefs = multiscale_flow_bp(pv, crfs, sws)
    torch.cuda.synchronize()
    b_in, ob, oe = None, {}, {}
    for lvl in (1, 0):
        N, _, H, W, K = shapes[lvl]
        p = probs[lvl].astype(np.float64)
        if b_in is not None:
            size = (2 * b_in.shape[2], 2 * b_in.shape[3], 2 * b_in.shape[4] - 1)
            up = np.stack([F.interpolate(torch.from_numpy(b_in[:, c])[:, None], size=size, mode='trilinear')[:, 0]
                           .numpy() for c in range(2)], axis=1)
            p = p + up / 2.0
        aff = np.zeros((N, 4, 3, H, W))
        aff[:, :, :2] = 0.1
        aff[:, :, 2] = 2.5
        pw = O.StereoPairwise(pack_stereo_jump(aff))
        ones = np.ones((N, 4, H, W))
        cu, cv = -p[:, 0], -p[:, 1]
        Mu, _ = O.lbp_forward_iters(cu, pw, ones, np.zeros((N, 4, H, W, K)), 3, 1)
        Mv, _ = O.lbp_forward_iters(cv, pw, ones, Mu, 3, 1)
        ob[lvl] = b_in = np.stack([O.beliefs(cu, Mu), O.beliefs(cv, Mv)], axis=1)
        oe[lvl] = np.stack([cu + Mu.sum(axis=1), cv + Mv.sum(axis=1)], axis=1)
    for lvl in (0, 1):   # level 1: u then v (2 solves in a chain); level 0: 4 with the prior
        P.elementwise(f"beliefs level {lvl}", beliefs[lvl].cpu().numpy(), ob[lvl], tol=P.beliefs_bar(4 - 2 * lvl))
    dev = P.energy_deviation(beliefs[0].cpu().numpy(), oe[0])
    clear = P.winner_clear(oe[0], dev)
    print(f"    WTA: energy deviation up to {dev.max():.3g}, {1 - clear.mean():.2e} of pixels within twice their "
          f"own deviation of a tie", flush=True)
    assert clear.mean() > 0.99
    f0 = flows[0].cpu().numpy()
    assert np.array_equal(f0[clear], (ob[0].argmax(-1) - sws[0] // 2)[clear])
