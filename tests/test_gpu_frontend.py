"""GPU test of the semantic CRF model (semantic_segmentation.SemanticNet): the front-end
networks (PWNet / EdgeNet, seeded random weights) feed the semantic LBP kernels, and the
beliefs and labels must equal the FP64 oracle run on the same unaries, pairwise costs and
edge weights.  This pins the glue between the networks and the op (the [1, 2*C*C, H, W] ->
[2, H, W, C, C] reshape, the per-pixel shift of the pairwise, the edge-weight adjustment, the
softmax of the unaries and the winner-take-all) to the oracle."""
import types

import numpy as np
import pytest
import torch

import _parity as P
from oracle import lbp as O
from _lbp_cases import pack_semantic_global

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _shift_pixel(j2):
    """[2,H,W,C,C] -> [4,H,W,C,C] as message_passing_op_pw_pixel.py:27-38."""
    _, H, W, C, _ = j2.shape
    p = np.zeros((4, H, W, C, C))
    p[0] = j2[0]
    p[1, :, 1:] = np.swapaxes(j2[0, :, :-1], -1, -2)
    p[2] = j2[1]
    p[3, 1:] = np.swapaxes(j2[1, :-1], -1, -2)
    return p


def _adjust_weights(w2):
    """[N,2,H,W] -> [N,4,H,W] as bp_op_cuda.py:87-95."""
    N, _, H, W = w2.shape
    w = np.zeros((N, 4, H, W))
    w[:, 0] = w2[:, 0]
    w[:, 1, :, 1:] = w2[:, 0, :, :-1]
    w[:, 2] = w2[:, 1]
    w[:, 3, 1:, :] = w2[:, 1, :-1, :]
    return w


@pytest.mark.parametrize("kind", ["pixel", "global"])
def test_semantic_net_matches_oracle(kind):
    from bplayers_amd.semantic_segmentation import SemanticNet
    print(f"semantic net {kind}", flush=True)
    torch.manual_seed(3)
    C, H, W = 20, 13, 22
    args = types.SimpleNamespace(num_labels=C, num_bp_layers=1, activation='leakyrelu', with_bn=True,
                                 with_upconv=False, with_output_bn=False, multi_level_output=False, pad=(0, 0),
                                 pairwise_type=kind, with_edges=kind == "global")
    net = SemanticNet(DEV, args)
    rng = np.random.default_rng(4)
    if kind == "global":
        m = np.stack([rng.uniform(0, 1, (C, C)), rng.uniform(0, 1, (C, C))])[None]
        net._pw_mat.data = torch.tensor(m, dtype=torch.float32, device=DEV)
    img = torch.randn(1, 3, H, W, device=DEV)
    logits = 2.0 * torch.randn(1, C, H, W, device=DEV)
    with torch.no_grad():
        labels, beliefs, unary_only = net(img, logits)
        if kind == "pixel":
            pw = O.make_pairwise("semantic_pixel", _shift_pixel(net.pairwise_from_net(img).double().cpu().numpy()))
            edge = np.ones((1, 4, H, W))
        else:
            pw = O.make_pairwise("semantic_global", pack_semantic_global(m))
            edge = _adjust_weights(net._edge_net(img)[0].double().cpu().numpy())
    torch.cuda.synchronize()
    T_in = float(net._softmax_input.T.detach())
    lg = logits.double().cpu().numpy().transpose(0, 2, 3, 1) / T_in
    prob = np.exp(lg - lg.max(-1, keepdims=True))
    prob /= prob.sum(-1, keepdims=True)
    cost = -prob
    M, _ = O.lbp_forward_iters(cost, pw, edge, np.zeros((1, 4, H, W, C)), C, 1)
    b = O.beliefs(cost, M, float(net._crf[0].message_passing.softmin.T.detach()))
    bg = beliefs.double().cpu().numpy()
    assert bg.shape == (1, H, W, C)
    P.elementwise("beliefs", bg, b)
    # winner take all: identical where the oracle's energy minimum leads the runner-up by more
    # than twice the GPU's largest energy deviation (log-beliefs up to a constant)
    Tb = float(net._crf[0].message_passing.softmin.T.detach())
    E = (cost + M.sum(axis=1)) / Tb
    dev = P.energy_deviation(bg, E)
    clear = P.winner_clear(E, dev)
    print(f"    WTA: energy deviation up to {dev.max():.3g}, {1 - clear.mean():.2e} of pixels within twice their "
          f"own deviation of a tie", flush=True)
    lab = labels.cpu().numpy().reshape(1, H, W)
    assert clear.mean() > 0.99
    assert np.array_equal(lab[clear], b.argmax(-1)[clear])
    sp = np.sort(prob[0], axis=-1)
    uclear = (sp[..., -1] - sp[..., -2]) > 1e-5
    assert np.array_equal(unary_only.cpu().numpy().reshape(H, W)[uclear], prob[0].argmax(-1)[uclear])
