"""GPU parity of the LBP backward against the FP64 oracle: the exact adjoint (mode E) and
the reference's backward (mode R), for every pairwise variant and label layout."""
import numpy as np
import pytest
import torch

import _parity as P
from oracle import lbp as O
from test_gpu_lbp_fwd import make_case, _mods, check_argmins

pytestmark = pytest.mark.gpu


def oracle_backward(kind, mode, cost, jump, edge, G, Mg, lab, fg, sg, absolute=False):
    """The FP64 backward on the GPU forward's own outputs (messages, argmins, scales): both
    sides replay the same routing, as the reference's backward_minsum replays its inputs."""
    pw = O.make_pairwise(kind, jump)
    if mode == "exact":
        return O.lbp_backward_exact(cost, pw, edge, Mg, lab, fg, G, absolute=absolute)
    return O.lbp_backward_reference(cost, pw, edge, G, Mg, lab, sg, absolute=absolute)


def run_case(kind, N, H, W, C, delta, mode, seed, pass_delta=False, **extra):
    cost, jump, edge, M0 = make_case(kind, N, H, W, C, delta, seed=seed, **extra)
    mod = _mods()[kind]
    dev = torch.device("cuda:0")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    rng = np.random.default_rng(seed + 1)
    G = rng.normal(size=M0.shape).astype(np.float32)
    mo = torch.empty_like(t(M0))
    msgs, am, sc = mod.forward_minsum(t(cost), t(jump), t(edge), t(M0), delta, out=mo)
    if kind == "semantic_pixel":
        res = mod.backward_minsum(t(cost), t(edge), t(G), msgs, am, sc, mode=mode,
                                  jump=t(jump) if mode == "exact" else None)
    else:
        kw = {"delta": delta} if pass_delta else {}
        res = mod.backward_minsum(t(cost), t(jump), t(edge), t(G), msgs, am, sc, mode=mode, **kw)
    torch.cuda.synchronize()
    gc, gj, ge, gm = [r.cpu().numpy().astype(np.float64) for r in res]
    from bplayers_amd._lbp_op import _gargmin_of
    Mg = msgs.cpu().numpy().astype(np.float64)
    ag, gg = am.cpu().numpy(), _gargmin_of(sc).cpu().numpy()
    sg = sc.cpu().numpy().astype(np.float64)
    fg = ag == 255
    lab = np.where(fg, gg[..., None], ag).astype(np.int64)
    # the routing replayed here is the GPU forward's; check it against the FP64 step replay
    check_argmins(kind, cost, jump, edge, M0, delta, Mg, ag, gg)
    ora = oracle_backward(kind, mode, cost, jump, edge, G, Mg, lab, fg, sg)
    if kind == "semantic_global":
        assert gj.shape[0] == N and np.all(gj[1:] == 0), "batch entries past 0 of the matrix gradient"
        gj = gj[:1]
    for name, g, o in zip(("grad cost", "grad pairwise", "grad edge", "grad messages"), (gc, gj, ge, gm), ora):
        P.normwise(name, g, o)
    return True


CASES = [
    ("stereo", 2, 7, 9, 64, 3, {}),
    ("stereo", 1, 5, 6, 192, 3, {}),
    ("stereo", 1, 4, 7, 100, 3, {}),
    ("stereo", 2, 6, 8, 25, 3, {}),
    ("stereo", 1, 5, 6, 9, 3, {}),
    ("stereo", 1, 5, 6, 12, 12, {}),
    ("stereo", 1, 5, 7, 40, 3, {"K_in": 5}),
    ("stereo", 1, 5, 7, 30, 2, {"ordered": True}),
    ("stereo", 1, 1, 8, 16, 3, {}),
    ("stereo", 1, 8, 1, 16, 3, {}),
    # longer rows: the paired horizontal adjoint (exact mode) with several groups in each half
    ("stereo", 1, 3, 37, 64, 3, {}),
    ("stereo", 2, 2, 23, 100, 3, {}),
    ("stereo", 1, 2, 41, 192, 3, {}),
    ("stereo", 1, 4, 13, 64, 3, {"offset": True}),
    ("stereo", 1, 3, 2, 64, 3, {}),
    # paired workgroups hold two rows: one row only, and a batch of two with an odd row count
    ("stereo", 1, 1, 9, 64, 3, {}),
    ("stereo", 2, 3, 11, 192, 3, {}),
    # 192 < C <= 255: four labels per lane (the LPL = 4 paired and unpaired kernels)
    ("stereo", 1, 4, 5, 255, 3, {}),
    ("stereo", 2, 3, 11, 200, 3, {}),
    # one pairwise tensor shared by a batch of two: per-image launches summing grad_jump
    ("stereo", 2, 4, 9, 64, 3, {"shared_jump": True}),
    ("stereo", 2, 3, 7, 192, 3, {"shared_jump": True}),
    ("semantic_global", 2, 5, 6, 19, 19, {}),
    ("semantic_global", 1, 4, 5, 19, 3, {}),
    ("semantic_global", 1, 3, 30, 25, 25, {}),
    ("semantic_global", 2, 7, 5, 32, 32, {}),
    ("semantic_pixel", 1, 4, 5, 7, 7, {}),
    ("semantic_pixel", 2, 4, 5, 19, 19, {}),
    ("semantic_pixel", 2, 5, 7, 19, 19, {}),
    ("semantic_pixel", 2, 3, 9, 25, 25, {}),
    ("semantic_pixel", 1, 4, 5, 32, 32, {}),
    ("semantic_pixel", 1, 5, 6, 21, 21, {}),
    ("semantic_pixel", 2, 4, 6, 19, 3, {}),
    ("semantic_pixel", 3, 3, 4, 7, 7, {}),
    # single rows / columns: passes with no step
    ("semantic_global", 1, 1, 8, 19, 19, {}),
    ("semantic_global", 2, 6, 1, 25, 25, {}),
    ("semantic_pixel", 2, 7, 1, 19, 19, {}),
    ("semantic_pixel", 1, 1, 6, 21, 21, {}),
]


# the stereo and semantic-global backward calls take the forward's delta (the semantic-pixel
# module's has no such argument): with it, delta >= C tells the semantic-global kernel that no
# argmin marks a truncation-term winner
RUNS = [(c, pd) for c in CASES for pd in ((False, True) if c[0] in ("stereo", "semantic_global") else (False,))]


@pytest.mark.parametrize("case,pass_delta", RUNS,
                         ids=lambda v: f"{v[0]}-N{v[1]}H{v[2]}W{v[3]}C{v[4]}d{v[5]}{''.join(v[6])}"
                         if isinstance(v, tuple) else ("delta" if v else "no-delta"))
@pytest.mark.parametrize("mode", ["exact", "reference"])
def test_lbp_backward_matches_oracle(case, mode, pass_delta):
    """'delta' passes the forward's delta, which selects the stereo fast step where it applies
    (delta <= 3, C > 32) and the semantic-global step without truncation-term winners
    (delta >= C); 'no-delta' runs the general kernels."""
    kind, N, H, W, C, delta, extra = case
    print(f"case {case} mode={mode} pass_delta={pass_delta}", flush=True)
    run_case(kind, N, H, W, C, delta, mode, seed=100 + C + H, pass_delta=pass_delta, **extra)


@pytest.mark.parametrize("kind,C", [("stereo", 192), ("stereo", 64), ("semantic_global", 19), ("semantic_pixel", 19)])
def test_reference_mode_passes_in_grad_through(kind, C):
    """Reference mode with grad_messages_out = in_grad: no copy is made, in_grad is left as it
    was, and every other gradient equals the copying call's bit for bit."""
    cost, jump, edge, M0 = make_case(kind, 1, 5, 13, C, 3 if kind == "stereo" else C, seed=31)
    mod = _mods()[kind]
    dev = torch.device("cuda:0")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    cost, jump, edge = t(cost), t(jump), t(edge)
    msgs, am, sc = mod.forward_minsum(cost, jump, edge, t(M0), 3 if kind == "stereo" else C, out=torch.empty_like(t(M0)))
    G = t(np.random.default_rng(3).normal(size=M0.shape).astype(np.float32))
    G0 = G.clone()
    if kind == "semantic_pixel":
        a = mod.backward_minsum(cost, edge, G, msgs, am, sc)
        b = mod.backward_minsum(cost, edge, G, msgs, am, sc, grad_messages_out=G)
    else:
        a = mod.backward_minsum(cost, jump, edge, G, msgs, am, sc)
        b = mod.backward_minsum(cost, jump, edge, G, msgs, am, sc, grad_messages_out=G)
    torch.cuda.synchronize()
    assert b[3].data_ptr() == G.data_ptr() and torch.equal(G, G0) and torch.equal(a[3], G0)
    for x, y in zip(a[:3], b[:3]):
        assert torch.equal(x, y)
    with pytest.raises(RuntimeError):
        if kind == "semantic_pixel":
            mod.backward_minsum(cost, edge, G, msgs, am, sc, mode="exact", jump=jump, grad_messages_out=G)
        else:
            mod.backward_minsum(cost, jump, edge, G, msgs, am, sc, mode="exact", grad_messages_out=G)



@pytest.mark.parametrize("kind,N,H,W,C,jn,mode", [
    ("stereo", 1, 6, 40, 192, None, "exact"),      # paired horizontal kernel (fast step)
    ("stereo", 2, 5, 9, 100, 1, "exact"),          # one jump for the batch: per-image launches
    ("stereo", 1, 4, 7, 64, None, "reference"),    # unpaired fast kernels and the combine pass
    ("semantic_global", 2, 5, 6, 19, None, "exact"),
])
def test_accumulate_adds_to_the_outputs(kind, N, H, W, C, jn, mode):
    """BPL_GRAD_ACCUMULATE: grad_cost, grad_pairwise and grad_edge come out as the tensors
    passed in plus the plain call's results (bit for bit: one fp32 addition each); the
    message gradient is the plain call's."""
    from bplayers_amd import _lbp_op, _native as nat
    knd = {"stereo": nat.LBP_STEREO, "semantic_global": nat.LBP_SEMANTIC_GLOBAL}[kind]
    cost, jump, edge, M0 = make_case(kind, N, H, W, C, 3, seed=7 + C)
    if jn == 1:
        jump = jump[:1]
    dev = torch.device("cuda:0")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    rng = np.random.default_rng(C)
    G = t(rng.normal(size=M0.shape).astype(np.float32))
    msgs, am, sc = _lbp_op.forward_minsum(knd, t(cost), t(jump), t(edge), t(M0), 3, out=torch.empty_like(t(M0)))
    plain = _lbp_op.backward_minsum(knd, t(cost), t(jump), t(edge), G, msgs, am, sc, mode=mode, delta=3)
    prior = [torch.randn(r.shape, generator=torch.Generator().manual_seed(i)).to(dev) for i, r in enumerate(plain[:3])]
    if kind == "semantic_global":
        prior[1][1:] = 0   # the matrix gradient lives in batch entry 0
    expect = [p + r for p, r in zip(prior, plain[:3])]
    acc = _lbp_op.backward_minsum(knd, t(cost), t(jump), t(edge), G, msgs, am, sc, mode=mode, delta=3,
                                  accumulate_into=[p.clone() for p in prior])
    torch.cuda.synchronize()
    for name, a, e in zip(("grad cost", "grad pairwise", "grad edge"), acc[:3], expect):
        d = (a - e).abs().max().item()
        print(f"    {kind} {mode}: accumulated {name} max |diff| {d:.3g}", flush=True)
        assert torch.equal(a, e), f"accumulated {name} differs from prior + plain"
    assert torch.equal(acc[3], plain[3]), "message gradient changed by accumulation"


@pytest.mark.parametrize("C", [192, 255])
def test_chain_accumulates_like_explicit_sums(C):
    """exact_chain_backward (accumulating in the kernels, chained row-end flags) against the
    same three calls without accumulation summed in torch, in the same order."""
    from bplayers_amd import _lbp_op, _native as nat
    N, H, W, iters = 1, 6, 37, 3
    cost, jump, edge, M0 = make_case("stereo", N, H, W, C, 3, seed=11)
    dev = torch.device("cuda:0")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    G = t(np.random.default_rng(3).normal(size=M0.shape).astype(np.float32))
    _, recs = _lbp_op.exact_chain_forward(nat.LBP_STEREO, t(cost), t(jump), t(edge), t(M0), 3, iters)
    got = _lbp_op.exact_chain_backward(nat.LBP_STEREO, t(cost), t(jump), t(edge), G, recs, delta=3)
    g, tot = G, None
    for it in reversed(range(iters)):
        m, am, sc = recs[it]
        gc, gj, ge, g = _lbp_op.backward_minsum(nat.LBP_STEREO, t(cost), t(jump), t(edge), g, m, am, sc,
                                                mode="exact", delta=3, in_h_row_ends=it < iters - 1,
                                                out_h_row_ends=it > 0)
        tot = [gc, gj, ge] if tot is None else [tot[0] + gc, tot[1] + gj, tot[2] + ge]
    torch.cuda.synchronize()
    for name, a, e in zip(("grad cost", "grad pairwise", "grad edge"), got[:3], tot):
        print(f"    C={C} chain {name}: max |diff| {(a - e).abs().max().item():.3g}", flush=True)
        assert torch.equal(a, e), f"chain {name}: in-kernel accumulation differs from the explicit sum"
    assert torch.equal(got[3], g)
