"""GPU parity of the LBP forward (libbpl_amd via bplayers_amd) against the FP64 oracle."""
import numpy as np
import pytest
import torch

import _parity as P
from oracle import lbp as O
from _lbp_cases import random_stereo_problem, semantic_matrix, pack_semantic_global, LEFT, RIGHT, UP, DOWN

pytestmark = pytest.mark.gpu


def _mods():
    from bplayers_amd import pytorch_cuda_lbp_op_stereo as st
    from bplayers_amd import pytorch_cuda_lbp_op_semantic_pw as sg
    from bplayers_amd import pytorch_cuda_lbp_op_semantic_pw_pixel as sp
    return {"stereo": st, "semantic_global": sg, "semantic_pixel": sp}


def make_case(kind, N, H, W, C, delta, seed, K_in=3, offset=False, ordered=False, shared_jump=False):
    rng = np.random.default_rng(seed)
    if kind == "stereo":
        cost, aff, jump, edge = random_stereo_problem(rng, N, H, W, C, K_in=K_in, ordered=ordered)
        if offset:
            jump[:, :, 0] = rng.choice([0.0, 0.5, -1.0, 1.3, 2.5], size=jump[:, :, 0].shape)
        if shared_jump:  # one pairwise tensor [1,4,K,H,W] for the whole batch
            jump = jump[:1]
    elif kind == "semantic_global":
        cost = rng.normal(size=(N, H, W, C))
        jump = pack_semantic_global(np.stack([semantic_matrix(rng, C), semantic_matrix(rng, C)])[None])
        edge = rng.uniform(0.5, 1.5, size=(N, 4, H, W))
    else:
        cost = rng.normal(size=(N, H, W, C))
        jump = rng.uniform(0, 1, size=(4, H, W, C, C))
        edge = rng.uniform(0.5, 1.5, size=(N, 4, H, W))
    M0 = rng.normal(size=(N, 4, H, W, C)) * 0.3
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)
    return f32(cost), f32(jump), f32(edge), f32(M0)


def run_gpu(kind, cost, jump, edge, M0, delta, inplace=True):
    from bplayers_amd._lbp_op import _gargmin_of
    mod = _mods()[kind]
    dev = torch.device("cuda:0")
    t = lambda a: torch.from_numpy(a).to(dev)
    m = t(M0)
    out = None if inplace else torch.empty_like(m)
    msgs, am, sc = mod.forward_minsum(t(cost), t(jump), t(edge), m, delta, out=out)
    torch.cuda.synchronize()
    if inplace:
        assert msgs.data_ptr() == m.data_ptr()
    else:
        assert torch.equal(m.cpu(), torch.from_numpy(M0)), "input messages modified"
    return (msgs.cpu().numpy().astype(np.float64), am.cpu().numpy(), sc.cpu().numpy().astype(np.float64),
            _gargmin_of(sc).cpu().numpy())


def written_mask(N, H, W):
    """Receiver slots a forward pass writes (the first pixel of each pass receives nothing)."""
    w = np.ones((N, 4, H, W), dtype=bool)
    w[:, LEFT, :, 0] = w[:, RIGHT, :, -1] = False
    w[:, UP, 0, :] = w[:, DOWN, -1, :] = False
    return w


def check_argmins(kind, cost, jump, edge, M0, delta, Mg, ag, gg, oracle=None):
    """GPU argmins against (i) the FP64 replay of every step from the GPU's own stored inputs,
    under the fp32 rounding-bound tie mask, and (ii) the FP64 oracle's own chain, at the
    decisions whose oracle margin exceeds what the GPU's input deviation can move."""
    pw = O.make_pairwise(kind, jump)
    gpu_fg = ag == 255
    gpu_lab = np.where(gpu_fg, np.broadcast_to(gg[..., None], ag.shape), ag).astype(np.int64)
    lab, fg, clear, written = O.local_argmins(cost, pw, edge, M0, Mg.astype(np.float32), delta)
    P.argmins("argmins vs FP64 step replay", gpu_lab, gpu_fg, lab, fg, clear, written)
    if oracle is not None:
        Mo, ao, fgo, lso, margin = oracle
        # a decision of the oracle's chain can only flip if its margin is within the GPU's
        # deviation of the step's inputs (chain + two cross messages) plus rounding
        dev = np.abs(Mg - Mo).max()
        thr = 3.0 * dev + 2.0 * O.ROUND_KAPPA * O.U32 * (np.abs(cost).max() + 3 * np.abs(Mo).max() + 10.0)
        ora_lab = np.where(fgo, np.broadcast_to(lso[..., None], ao.shape), ao)
        P.argmins("argmins vs FP64 oracle chain", gpu_lab, gpu_fg, ora_lab, fgo, margin > thr, written)


def check_forward(kind, cost, jump, edge, M0, delta, inplace=True):
    Mg, ag, sg, gg = run_gpu(kind, cost, jump, edge, M0, delta, inplace)
    pw = O.make_pairwise(kind, jump)
    Mo, ao, so, fgo, lso, margin = O.lbp_forward(cost, pw, edge, M0, delta, want_margin=True)
    N, _, H, W, C = Mo.shape
    written = written_mask(N, H, W)
    err = P.rows("messages", Mg[written], Mo[written])
    P.elementwise("message scale", sg[written], so[written])
    check_argmins(kind, cost, jump, edge, M0, delta, Mg, ag, gg, (Mo, ao, fgo, lso, margin))
    # unwritten boundary slots are copies of the input
    np.testing.assert_array_equal(Mg[~written], M0.astype(np.float64)[~written])
    return err


CASES = [
    # kind, N, H, W, C, delta, extra
    ("stereo", 2, 9, 13, 64, 3, {}),
    ("stereo", 1, 6, 7, 192, 3, {}),
    ("stereo", 1, 5, 9, 100, 3, {}),
    ("stereo", 1, 4, 5, 255, 3, {}),
    ("stereo", 2, 7, 11, 25, 3, {}),
    ("stereo", 1, 6, 6, 9, 3, {}),
    ("stereo", 1, 5, 6, 12, 12, {}),
    ("stereo", 1, 6, 7, 40, 3, {"K_in": 5}),
    ("stereo", 1, 6, 7, 40, 3, {"offset": True}),
    ("stereo", 1, 6, 7, 30, 2, {"ordered": True}),
    ("stereo", 1, 1, 9, 16, 3, {}),
    ("stereo", 1, 9, 1, 16, 3, {}),
    ("semantic_global", 2, 6, 7, 19, 19, {}),
    ("semantic_global", 1, 5, 6, 19, 3, {}),
    ("semantic_global", 1, 4, 5, 70, 70, {}),
    ("semantic_global", 1, 3, 40, 25, 25, {}),
    ("semantic_global", 2, 9, 6, 32, 32, {}),
    ("semantic_global", 1, 5, 7, 7, 7, {}),
    ("semantic_pixel", 1, 5, 6, 7, 7, {}),
    ("semantic_pixel", 1, 4, 5, 19, 19, {}),
    ("semantic_pixel", 2, 5, 7, 19, 19, {}),
    ("semantic_pixel", 2, 3, 9, 25, 25, {}),
    ("semantic_pixel", 1, 4, 5, 32, 32, {}),
    ("semantic_pixel", 1, 5, 6, 19, 3, {}),
    ("semantic_pixel", 3, 3, 4, 7, 7, {}),
    # single rows / columns: passes with no step
    ("semantic_global", 1, 1, 8, 19, 19, {}),
    ("semantic_global", 2, 6, 1, 25, 25, {}),
    ("semantic_pixel", 2, 7, 1, 19, 19, {}),
    ("semantic_pixel", 1, 1, 6, 21, 21, {}),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}-N{c[1]}H{c[2]}W{c[3]}C{c[4]}d{c[5]}{''.join(c[6])}")
@pytest.mark.parametrize("inplace", [True, False])
@pytest.mark.parametrize("impl", ["default", "generic"])
def test_lbp_forward_matches_oracle(case, inplace, impl, monkeypatch):
    """'default' runs the specialised kernels where they apply (stereo 64-lane kernels, semantic
    full min-plus kernels); 'generic' forces the generic kernels (BPL_LBP_GENERIC=1)."""
    kind, N, H, W, C, delta, extra = case
    if impl == "generic":
        monkeypatch.setenv("BPL_LBP_GENERIC", "1")
    print(f"case {case} inplace={inplace} impl={impl}", flush=True)
    cost, jump, edge, M0 = make_case(kind, N, H, W, C, delta, seed=hash((N, H, W, C)) % 1000, **extra)
    check_forward(kind, cost, jump, edge, M0, delta, inplace)


@pytest.mark.parametrize("C", [64, 192])
def test_lbp_forward_two_iterations_long_rows(C):
    """Chains long enough for fp32 drift to show (W=300)."""
    cost, jump, edge, M0 = make_case("stereo", 1, 4, 300, C, 3, seed=5)
    from bplayers_amd import pytorch_cuda_lbp_op_stereo as st
    dev = torch.device("cuda:0")
    t = lambda a: torch.from_numpy(a).to(dev)
    m = t(M0)
    for _ in range(2):
        st.forward_minsum(t(cost), t(jump), t(edge), m, 3)
    M = M0.astype(np.float64)
    for _ in range(2):
        M, *_ = O.lbp_forward(cost, O.StereoPairwise(jump), edge, M, 3)
    P.rows("messages after 2 iterations, W=300", m.cpu().numpy(), M)


@pytest.mark.parametrize("C", [40, 64, 100, 192])
@pytest.mark.parametrize("offset", [False, True])
def test_lbp_forward_messages_only(C, offset):
    """Forward without argmins and scales (the serving call): the same messages as the full
    call, bit for bit."""
    from bplayers_amd import _native as nat
    from bplayers_amd import pytorch_cuda_lbp_op_stereo as st
    print(f"messages-only C={C} offset={offset}", flush=True)
    cost, jump, edge, M0 = make_case("stereo", 1, 7, 37, C, 3, seed=C, offset=offset)
    dev = torch.device("cuda:0")
    t = lambda a: torch.from_numpy(a).to(dev)
    full, _, _ = st.forward_minsum(t(cost), t(jump), t(edge), t(M0), 3, out=torch.empty_like(t(M0)))
    c_, j_, e_, m_ = t(cost), t(jump), t(edge), t(M0)
    out = torch.empty_like(m_)
    N, H, W = 1, 7, 37
    nat.call("bpl_lbp_fwd", nat.LBP_STEREO, nat.ptr(c_), nat.ptr(j_), j_.shape[2], 1, nat.ptr(e_), nat.ptr(m_),
             nat.ptr(out), None, None, None, N, H, W, C, 3, nat.stream_of(c_))
    torch.cuda.synchronize()
    assert torch.equal(out, full)


BASE_CASES = [  # (N, H, W, C, shared jump, argmins, in place, break): bpl_lbp_fwd_ws vs bpl_lbp_fwd
    (1, 5, 40, 64, False, True, False, None),        # one label per lane, second visits present
    (2, 3, 41, 100, True, True, True, None),         # partial lanes, odd row count, shared jump, in place
    (1, 4, 9, 192, False, True, False, None),        # rows too short for second visits
    (1, 6, 57, 255, False, False, False, None),      # four labels per lane, messages only
    (2, 5, 40, 192, False, True, False, "offset"),   # a pixel with a window offset: unpaired rows
    (1, 4, 40, 64, False, True, False, "big"),       # a scanline whose truncation term can underflow
    (1, 8, 128, 192, False, True, False, None),      # several ring groups of second visits
]


@pytest.mark.parametrize("case", BASE_CASES, ids=[f"{c[0]}x{c[1]}x{c[2]}x{c[3]}-{c[7]}" for c in BASE_CASES])
def test_lbp_forward_base_plane_bitwise(case):
    """The workspace form (horizontal phase writes (cost + M_R) + M_L, vertical P1/P2 sweeps read
    it) gives the outputs of bpl_lbp_fwd bit for bit, paired and unpaired rows alike."""
    from bplayers_amd import _native as nat
    from bplayers_amd._lbp_op import _alloc_scale_and_gargmin
    N, H, W, C, shared, outs, inplace, brk = case
    print(f"base plane case {case}", flush=True)
    cost, jump, edge, M0 = make_case("stereo", N, H, W, C, 3, seed=W + C, shared_jump=shared)
    if brk == "offset":
        jump[0, RIGHT, 0, 2, 17] = 1.0
    elif brk == "big":
        edge[0, LEFT, 1, :] = 1.0
        edge[0, LEFT, 1, 20] = 200.0
    dev = torch.device("cuda:0")
    t = lambda a: torch.from_numpy(a).to(dev)
    c_, j_, e_ = t(cost), t(jump), t(edge)
    K, Nj = j_.shape[2], j_.shape[0]
    res = []
    for use_ws in (False, True):
        m_in = t(M0)
        out = m_in if inplace else torch.empty_like(m_in)
        am = torch.empty((N, 4, H, W, C), dtype=torch.uint8, device=dev) if outs else None
        sc, ga = _alloc_scale_and_gargmin(N, H, W, dev) if outs else (None, None)
        wsb = int(nat.lib().bpl_lbp_fwd_workspace_size(nat.LBP_STEREO, N, H, W, C))
        assert wsb == N * H * W * C * 4
        ws = torch.full((wsb // 4,), float("nan"), device=dev) if use_ws else None
        args = [nat.LBP_STEREO, nat.ptr(c_), nat.ptr(j_), K, Nj, nat.ptr(e_), nat.ptr(m_in), nat.ptr(out),
                nat.ptr(am), nat.ptr(ga), nat.ptr(sc), N, H, W, C, 3]
        if use_ws:
            nat.call("bpl_lbp_fwd_ws", *args, nat.ptr(ws), wsb, 0, nat.stream_of(c_))
        else:
            nat.call("bpl_lbp_fwd", *args, nat.stream_of(c_))
        torch.cuda.synchronize()
        res.append((out.clone(), am, sc, ga, ws))
    (m0, a0, s0, g0, _), (m1, a1, s1, g1, ws) = res
    assert torch.equal(m0, m1), "messages differ"
    if outs:
        assert torch.equal(a0, a1) and torch.equal(s0, s1) and torch.equal(g0, g1), "argmins or scales differ"
    # every base entry was written: (cost + M_R) + M_L of the output messages
    base = ws.view(N, H, W, C)
    ref = (c_ + m1[:, RIGHT]) + m1[:, LEFT]
    assert torch.equal(base, ref), "base plane differs from (cost + M_R) + M_L"


def test_lbp_forward_ws_rejects_overlap():
    from bplayers_amd import _native as nat
    N, H, W, C = 1, 3, 8, 64
    dev = torch.device("cuda:0")
    cost = torch.zeros((N, H, W, C), device=dev)
    m = torch.zeros((N, 4, H, W, C), device=dev)
    jump = torch.zeros((N, 4, 6, H, W), device=dev)
    edge = torch.ones((N, 4, H, W), device=dev)
    wsb = N * H * W * C * 4
    with pytest.raises(nat.BplError, match="overlaps"):
        nat.call("bpl_lbp_fwd_ws", nat.LBP_STEREO, nat.ptr(cost), nat.ptr(jump), 6, 1, nat.ptr(edge), nat.ptr(m),
                 nat.ptr(m), None, None, None, N, H, W, C, 3, nat.ptr(m), wsb, 0, nat.stream_of(cost))
    with pytest.raises(nat.BplError, match="needed"):
        nat.call("bpl_lbp_fwd_ws", nat.LBP_STEREO, nat.ptr(cost), nat.ptr(jump), 6, 1, nat.ptr(edge), nat.ptr(m),
                 nat.ptr(m), None, None, None, N, H, W, C, 3, nat.ptr(torch.empty(4, device=dev)), 16, 0,
                 nat.stream_of(cost))


@pytest.mark.parametrize("case", [(1, 5, 40, 192, True), (2, 3, 41, 100, False), (1, 4, 9, 64, True)])
def test_lbp_forward_zero_input_messages(case):
    """BPL_FWD_MSG_IN_ZERO (the first iteration from zero messages, whose planes the P1/P2 sweeps
    then do not read): the outputs of bpl_lbp_fwd on the same zero messages, bit for bit, with and
    without the base-plane workspace."""
    from bplayers_amd import _native as nat
    from bplayers_amd._lbp_op import _alloc_scale_and_gargmin
    N, H, W, C, inplace = case
    print(f"zero-input case {case}", flush=True)
    cost, jump, edge, _ = make_case("stereo", N, H, W, C, 3, seed=7 + W)
    dev = torch.device("cuda:0")
    t = lambda a: torch.from_numpy(a).to(dev)
    c_, j_, e_ = t(cost), t(jump), t(edge)
    res = []
    for mode in ("plain", "flag", "flag+ws"):
        m_in = torch.zeros((N, 4, H, W, C), device=dev)
        out = m_in if inplace else torch.empty_like(m_in)
        am = torch.empty((N, 4, H, W, C), dtype=torch.uint8, device=dev)
        sc, ga = _alloc_scale_and_gargmin(N, H, W, dev)
        args = [nat.LBP_STEREO, nat.ptr(c_), nat.ptr(j_), 6, N, nat.ptr(e_), nat.ptr(m_in), nat.ptr(out),
                nat.ptr(am), nat.ptr(ga), nat.ptr(sc), N, H, W, C, 3]
        if mode == "plain":
            nat.call("bpl_lbp_fwd", *args, nat.stream_of(c_))
        else:
            ws = torch.empty(N * H * W * C, device=dev) if mode == "flag+ws" else None
            nat.call("bpl_lbp_fwd_ws", *args, nat.ptr(ws), 0 if ws is None else ws.numel() * 4,
                     nat.FWD_MSG_IN_ZERO, nat.stream_of(c_))
        torch.cuda.synchronize()
        res.append((out.clone(), am, sc, ga))
    for r in res[1:]:
        assert all(torch.equal(a, b) for a, b in zip(res[0], r)), "outputs differ"
