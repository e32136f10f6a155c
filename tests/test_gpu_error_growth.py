"""How the fp32 kernels' error against FP64 grows with chain length and iterations, on the
stereo workload's own inputs (features warped by a smooth disparity field, cost =
-softmax(-SAD), per-pixel softplus P1 / P2), 4-row strips of widths 64 to 1248.

For every size and iteration 1..3 the test prints the GPU's error next to that of the C++
fp32 restatement of the reference (oracle/c, the reference's arithmetic in plain fp32), both
against the FP64 restatement, for the messages (per vector), the beliefs (elementwise) and the
exact backward's cost gradient (normwise).  It asserts that the GPU is as accurate as fp32
arithmetic gets here: within the bars of _parity (1e-5; beliefs_bar), or within 4x the plain
fp32 restatement's error (maxima over a million values differ by that between two fp32
evaluation orders)."""
import numpy as np
import pytest
import torch

import _parity as P
from oracle import cpu as OC
from oracle import lbp as O

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _inputs(H, W, D):
    from bplayers_amd import workloads
    from oracle import sad as S
    from _lbp_cases import pack_stereo_jump
    wl = workloads.stereo_full(H=H, W=W, D=D)
    sad = S.stereo_sad_forward(wl["f0"], wl["f1"], 0, D - 1)
    z = -sad
    p = np.exp(z - z.max(-1, keepdims=True))
    p /= p.sum(-1, keepdims=True)
    return (-p).astype(np.float32), pack_stereo_jump(wl["affinities"]).astype(np.float32)


def _bar(ref_err, floor=P.TOL):
    return max(floor, 4.0 * ref_err)


@pytest.mark.parametrize("D,W", [(192, 64), (192, 312), (192, 1248), (64, 1248)])
def test_error_growth_with_chain_length(D, W):
    from bplayers_amd import _lbp_op, _native as nat
    H, iters = 4, 3
    cost, jump = _inputs(H, W, D)
    N = 1
    edge = np.ones((N, 4, H, W), dtype=np.float32)
    print(f"error growth: H={H} W={W} D={D}", flush=True)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(DEV)
    _, recs = _lbp_op.exact_chain_forward(nat.LBP_STEREO, t(cost), t(jump), t(edge),
                                          torch.zeros((N, 4, H, W, D), device=DEV), 3, iters)
    M64 = np.zeros((N, 4, H, W, D))
    M32 = np.zeros((N, 4, H, W, D), dtype=np.float32)
    wr = np.ones((N, 4, H, W), dtype=bool)
    wr[:, 0, :, 0] = wr[:, 1, :, -1] = wr[:, 2, 0, :] = wr[:, 3, -1, :] = False
    rows = []
    for it in range(iters):
        Mg = recs[it][0].cpu().numpy().astype(np.float64)
        M64, _, _ = OC.lbp_forward(cost, jump, edge, M64, 3, dtype=np.float64)
        M32, _, _ = OC.lbp_forward(cost, jump, edge, M32, 3, dtype=np.float32)
        nw = lambda a: float((np.abs(a - M64)[wr].max(-1) / np.abs(M64[wr]).max(-1)).max())
        e_g, e_r = nw(Mg), nw(M32.astype(np.float64))
        E = cost + M64.sum(axis=1)
        b64 = O.softmin(E)
        m = b64 >= P.FP32_TINY
        be = lambda M: float((np.abs(O.softmin(cost + M.sum(axis=1)) - b64)[m] / b64[m]).max())
        # beliefs as the fp32 kernels would form them from their messages (FP64 softmin)
        b_g, b_r = be(Mg), be(M32.astype(np.float64))
        rows.append((it + 1, e_g, e_r, b_g, b_r))
        print(f"    iteration {it + 1}: messages GPU {e_g:.3g} / fp32 restatement {e_r:.3g}; "
              f"beliefs GPU {b_g:.3g} / fp32 restatement {b_r:.3g}", flush=True)
        assert e_g <= _bar(e_r), f"messages at iteration {it + 1}: {e_g:.3g} vs fp32 restatement {e_r:.3g}"
        assert b_g <= _bar(b_r, P.beliefs_bar(it + 1)), \
            f"beliefs at iteration {it + 1}: {b_g:.3g} vs fp32 restatement {b_r:.3g}"
    # exact backward of all iterations from the loss gradient of a weighted sum of the final
    # beliefs, on the GPU's forward state; against FP64 and the fp32 restatement on that state
    rng = np.random.default_rng(W + D)
    Wb = rng.normal(size=(N, H, W, D))
    Mf = recs[-1][0].cpu().numpy().astype(np.float64)
    gE, kE = P.loss_grad_of_beliefs(cost + Mf.sum(axis=1), Wb)
    G = np.ascontiguousarray(np.broadcast_to(gE[:, None], Mf.shape), dtype=np.float32)
    gc_g, _, _, _ = _lbp_op.exact_chain_backward(nat.LBP_STEREO, t(cost), t(jump), t(edge), t(G), recs, delta=3)
    gc_g = gc_g.cpu().numpy().astype(np.float64)
    states = []
    for m_, am, sc in recs:
        from bplayers_amd._lbp_op import _gargmin_of
        ag = am.cpu().numpy()
        fg = ag == 255
        lab = np.where(fg, _gargmin_of(sc).cpu().numpy()[..., None], ag).astype(np.int32)
        states.append((m_.cpu().numpy(), lab, fg.astype(np.uint8)))
    tot = {}
    for dt in (np.float64, np.float32):
        g = G.astype(dt)
        acc = 0.0
        for Mo, lab, fg in reversed(states):
            gc, _, _, g = OC.lbp_backward_exact(jump, edge, g, Mo, lab, fg, 3, dtype=dt)
            acc = acc + gc.astype(np.float64)
        tot[dt] = acc
    ref = tot[np.float64]
    ek = lambda a: float(np.abs(a - ref).max() / np.abs(ref).max())
    e_g, e_r = ek(gc_g), ek(tot[np.float32])
    print(f"    backward, {iters} iterations: grad cost normwise GPU {e_g:.3g} / fp32 restatement {e_r:.3g}",
          flush=True)
    assert e_g <= _bar(e_r)
