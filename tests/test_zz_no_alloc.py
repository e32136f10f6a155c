"""GPU checks of bpl_lbp_bwd_noalloc (the backward whose workspace is the library's built-in
scratch array): through the C ABI, its outputs equal bpl_lbp_bwd's with a caller-owned
workspace bit for bit, on two calls in a row (the second with a smaller input, so it runs on
what the first left in the scratch array), for each pairwise variant and both modes; sizes
whose workspace exceeds the array are rejected."""
import ctypes

import numpy as np
import pytest
import torch

from test_gpu_lbp_fwd import make_case, _mods

pytestmark = pytest.mark.gpu


def _grads(kind, jump, N, H, W, C, dev):
    from bplayers_amd import _native as nat
    gc = torch.empty((N, H, W, C), device=dev)
    gj = torch.empty((1, 4, C, C), device=dev) if kind == nat.LBP_SEMANTIC_GLOBAL else torch.empty_like(jump)
    ge = torch.empty((N, 4, H, W), device=dev)
    gm = torch.empty((N, 4, H, W, C), device=dev)
    for t in (gc, gj, ge, gm):
        t.fill_(float("nan"))   # every output must be written by both entry points
    return gc, gj, ge, gm


def _run(kind_name, N, H, W, C, delta, mode, seed):
    from bplayers_amd import _native as nat
    from bplayers_amd._lbp_op import _gargmin_of
    print(f"noalloc {kind_name} N{N} H{H} W{W} C{C} mode {mode}", flush=True)
    kind = {"stereo": nat.LBP_STEREO, "semantic_global": nat.LBP_SEMANTIC_GLOBAL,
            "semantic_pixel": nat.LBP_SEMANTIC_PIXEL}[kind_name]
    cost, jump, edge, M0 = make_case(kind_name, N, H, W, C, delta, seed=seed)
    dev = torch.device("cuda:0")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    cost, jump, edge = t(cost), t(jump), t(edge)
    msgs, am, sc = _mods()[kind_name].forward_minsum(cost, jump, edge, t(M0), delta, out=torch.empty_like(t(M0)))
    G = t(np.random.default_rng(seed).normal(size=M0.shape).astype(np.float32))
    K = jump.shape[2] if kind == nat.LBP_STEREO else 0
    Nj = jump.shape[0] if kind == nat.LBP_STEREO else 1
    garg = _gargmin_of(sc)
    m = nat.GRAD_EXACT if mode == "exact" else nat.GRAD_REFERENCE
    st = nat.stream_of(cost)
    p = nat.ptr
    ws_bytes = int(nat.lib().bpl_lbp_bwd_workspace_size(kind, N, H, W, C))
    assert ws_bytes <= nat.LBP_BWD_NOALLOC_MAX_BYTES
    ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=dev)
    a = _grads(kind, jump, N, H, W, C, dev)
    nat.call("bpl_lbp_bwd", kind, m, p(jump), K, Nj, p(edge), p(G), p(msgs), p(am), p(garg), *[p(x) for x in a],
             p(ws), ws_bytes, N, H, W, C, delta, st)
    b = _grads(kind, jump, N, H, W, C, dev)
    nat.call("bpl_lbp_bwd_noalloc", kind, m, p(jump), K, Nj, p(edge), p(G), p(msgs), p(am), p(garg),
             *[p(x) for x in b], N, H, W, C, delta, st)
    torch.cuda.synchronize()
    for name, x, y in zip(("grad cost", "grad pairwise", "grad edge", "grad messages"), a, b):
        assert not torch.isnan(x).any(), f"{name}: bpl_lbp_bwd left entries unwritten"
        assert torch.equal(x, y), f"{name}: the no-allocation entry point differs"


CASES = [
    # (kind, larger call, smaller call, delta, mode)
    ("stereo", (1, 6, 40, 192), (1, 3, 17, 100), 3, "exact"),       # paired fast kernels, then general rows
    ("stereo", (2, 5, 23, 64), (1, 4, 9, 64), 3, "reference"),
    ("semantic_global", (2, 6, 9, 19), (1, 4, 5, 19), 19, "exact"),  # C x C partials in the scratch array
    ("semantic_pixel", (2, 4, 6, 19), (1, 3, 5, 7), 19, "exact"),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}-{c[4]}")
def test_noalloc_backward_equals_workspace_backward(case):
    kind, big, small, delta, mode = case
    for i, (N, H, W, C) in enumerate((big, small)):
        _run(kind, N, H, W, C, min(delta, C), mode, seed=11 + i)


def test_noalloc_rejects_workspace_above_the_scratch_array():
    from bplayers_amd import _native as nat
    lib = nat.lib()
    N, H, W, C = 1, 1024, 1024, 64          # four [H,W,C] fp32 planes: 1 GiB
    assert lib.bpl_lbp_bwd_workspace_size(nat.LBP_STEREO, N, H, W, C) > nat.LBP_BWD_NOALLOC_MAX_BYTES
    P = ctypes.c_void_p(256)
    rc = lib.bpl_lbp_bwd_noalloc(nat.LBP_STEREO, nat.GRAD_EXACT, P, 6, 1, P, P, P, P, P, P, P, P, P, N, H, W, C,
                                 3, None)
    assert rc == -2 and b"scratch" in lib.bpl_last_error()
