"""GPU checks of the chained exact backward (in_h_row_ends / out_h_row_ends): a backward pass
through several iterations whose calls skip the LEFT/RIGHT slots of the message gradients away
from the row ends, and the DOWN slot away from rows 0 and H-1 (equal to the UP slot there),
gives bit-identical gradients to the same chain without the flags.

The skipped entries of every intermediate gradient are overwritten with NaN before the next
call reads it, so a kernel that still read one would put NaN into its outputs."""
import numpy as np
import pytest
import torch

from test_gpu_lbp_fwd import make_case, _mods

pytestmark = pytest.mark.gpu

LEFT, RIGHT, UP, DOWN = 0, 1, 2, 3
ITERS = 3


def _backward(mod, kind, cost, jump, edge, g, msgs, am, sc, delta, **flags):
    if kind == "semantic_pixel":
        return mod.backward_minsum(cost, edge, g, msgs, am, sc, mode="exact", jump=jump, **flags)
    if kind == "stereo":
        return mod.backward_minsum(cost, jump, edge, g, msgs, am, sc, mode="exact", delta=delta, **flags)
    return mod.backward_minsum(cost, jump, edge, g, msgs, am, sc, mode="exact", **flags)


def _poison_interior(gm):
    """NaN into the entries a chained gradient leaves unwritten: LEFT/RIGHT away from the row
    ends, DOWN away from rows 0 and H-1."""
    gm = gm.clone()
    gm[:, LEFT, :, 1:, :] = float("nan")
    gm[:, RIGHT, :, :-1, :] = float("nan")
    gm[:, DOWN, 1:-1] = float("nan")
    return gm


CASES = [
    # paired fast kernel (stereo, delta <= 3, K = 6), full and partial label rows
    ("stereo", 1, 6, 40, 192, 3, {}, True),
    ("stereo", 2, 5, 23, 100, 3, {}, True),
    ("stereo", 1, 4, 13, 64, 3, {"offset": True}, True),   # rows with offsets: general step
    ("stereo", 1, 5, 9, 64, 3, {}, False),                  # no delta: unpaired kernels + combine
    ("stereo", 1, 5, 7, 40, 3, {"K_in": 5}, True),          # K = 8
    ("stereo", 1, 5, 6, 12, 3, {}, True),                   # C <= 16: generic kernels
    ("stereo", 1, 3, 2, 64, 3, {}, True),                   # two-pixel rows
    ("stereo", 1, 1, 9, 64, 3, {}, True),                   # one row (paired workgroups hold two)
    ("stereo", 2, 3, 11, 192, 3, {}, True),                 # batch of two, odd row count
    ("semantic_global", 2, 5, 6, 19, 19, {}, False),
    ("semantic_global", 1, 4, 5, 12, 12, {}, False),
    ("semantic_pixel", 2, 4, 5, 19, 19, {}, False),
    ("semantic_pixel", 1, 4, 5, 7, 7, {}, False),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}-N{c[1]}H{c[2]}W{c[3]}C{c[4]}d{c[5]}"
                         f"{''.join(c[6])}{'-delta' if c[7] else ''}")
def test_chained_backward_equals_unchained(case):
    kind, N, H, W, C, delta, extra, pass_delta = case
    print(f"chain case {case}", flush=True)
    cost, jump, edge, M0 = make_case(kind, N, H, W, C, delta, seed=7 + C + W, **extra)
    mod = _mods()[kind]
    dev = torch.device("cuda:0")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    cost, jump, edge = t(cost), t(jump), t(edge)
    msgs, fwd = [t(M0)], []
    for _ in range(ITERS):
        out = torch.empty_like(msgs[-1])
        m, am, sc = mod.forward_minsum(cost, jump, edge, msgs[-1], delta, out=out)
        msgs.append(m)
        fwd.append((m, am, sc))
    G = torch.from_numpy(np.random.default_rng(11).normal(size=M0.shape).astype(np.float32)).to(dev)
    d = delta if pass_delta else None

    g_ref, g_chain = G, G
    for it in reversed(range(ITERS)):
        m, am, sc = fwd[it]
        ref = _backward(mod, kind, cost, jump, edge, g_ref, m, am, sc, d)
        chain = _backward(mod, kind, cost, jump, edge, g_chain, m, am, sc, d,
                          in_h_row_ends=it < ITERS - 1, out_h_row_ends=it > 0)
        torch.cuda.synchronize()
        for name, a, b in zip(("cost", "pairwise", "edge"), ref[:3], chain[:3]):
            assert torch.equal(a, b), f"iteration {it}: grad {name} differs"
        gr, gc = ref[3], chain[3]
        if it > 0:
            assert torch.equal(gr[:, UP], gc[:, UP]), f"iteration {it}: U message gradient differs"
            assert torch.equal(gr[:, DOWN, 0], gc[:, DOWN, 0]) and torch.equal(gr[:, DOWN, -1], gc[:, DOWN, -1])
            # the exact adjoint's D slot equals the U slot away from rows 0 and H-1
            assert torch.equal(gr[:, DOWN, 1:-1], gr[:, UP, 1:-1])
            assert torch.equal(gr[:, LEFT, :, 0], gc[:, LEFT, :, 0])
            assert torch.equal(gr[:, RIGHT, :, -1], gc[:, RIGHT, :, -1])
            # the exact adjoint's L/R slots really are zero away from the row ends
            assert torch.all(gr[:, LEFT, :, 1:] == 0) and torch.all(gr[:, RIGHT, :, :-1] == 0)
            g_chain = _poison_interior(gc)
        else:
            assert torch.equal(gr, gc), "bottom iteration: message gradient differs"
        g_ref = gr


KIND_CODES = {"stereo": 0, "semantic_global": 1, "semantic_pixel": 2}


@pytest.mark.parametrize("case", [c for c in CASES if c[0] != "semantic_pixel"] + [CASES[-2]],
                         ids=lambda c: f"{c[0]}-N{c[1]}H{c[2]}W{c[3]}C{c[4]}{'-delta' if c[7] else ''}")
def test_backward_without_message_gradient(case):
    """A null grad_msg (want_grad_messages=False: the bottom call of a chain from constant initial
    messages) leaves grad_cost, the pairwise and the edge gradients bitwise unchanged, in the
    exact mode (paired kernel, or combine pass) and the reference mode."""
    from bplayers_amd import _lbp_op
    kind, N, H, W, C, delta, extra, pass_delta = case
    print(f"no-message-gradient case {case}", flush=True)
    cost, jump, edge, M0 = make_case(kind, N, H, W, C, delta, seed=3 + C + W, **extra)
    mod = _mods()[kind]
    dev = torch.device("cuda:0")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    cost, jump, edge = t(cost), t(jump), t(edge)
    m, am, sc = mod.forward_minsum(cost, jump, edge, t(M0), delta, out=torch.empty_like(t(M0)))
    G = torch.from_numpy(np.random.default_rng(5).normal(size=M0.shape).astype(np.float32)).to(dev)
    d = delta if pass_delta else None
    for mode in ("exact", "reference"):
        if kind == "semantic_pixel" and mode == "reference":
            continue
        res = [_lbp_op.backward_minsum(KIND_CODES[kind], cost, jump, edge, G, m, am, sc, mode=mode, delta=d,
                                       want_grad_messages=w) for w in (True, False)]
        torch.cuda.synchronize()
        assert res[1][3] is None
        for name, a, b in zip(("cost", "pairwise", "edge"), res[0][:3], res[1][:3]):
            assert torch.equal(a, b), f"{mode}: grad {name} differs without the message gradient"
