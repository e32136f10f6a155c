"""GPU parity of the SAD ops against the FP64 oracle."""
import numpy as np
import pytest
import torch

import _parity as P
from oracle import sad as S

pytestmark = pytest.mark.gpu
U32 = 2.0 ** -24


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to("cuda:0")


@pytest.mark.parametrize("N,C,H,W,mn,mx", [(1, 32, 8, 100, 0, 63), (2, 64, 5, 130, 0, 191), (1, 3, 4, 17, -3, 5),
                                           (1, 20, 3, 70, 4, 70)])
def test_stereo_sad_forward(N, C, H, W, mn, mx):
    from bplayers_amd import pytorch_cuda_stereo_sad_op as op
    rng = np.random.default_rng(C + W)
    f0 = rng.normal(size=(N, C, H, W)).astype(np.float32)
    f1 = rng.normal(size=(N, C, H, W)).astype(np.float32)
    got = op.forward(_t(f0), _t(f1), mn, mx, 1.0).cpu().numpy()
    exp = S.stereo_sad_forward(f0, f1, mn, mx)
    P.elementwise("SAD volume", got, exp)


def test_stereo_sad_forward_fractional_step():
    from bplayers_amd import pytorch_cuda_stereo_sad_op as op
    rng = np.random.default_rng(1)
    f0 = rng.normal(size=(1, 4, 3, 20)).astype(np.float32)
    f1 = rng.normal(size=(1, 4, 3, 20)).astype(np.float32)
    got = op.forward(_t(f0), _t(f1), 0, 9, 0.5).cpu().numpy()
    P.elementwise("SAD volume, step 0.5", got, S.stereo_sad_forward(f0, f1, 0, 9, 0.5))


@pytest.mark.parametrize("N,C,H,W,mn,mx", [(1, 16, 4, 80, 0, 31), (2, 8, 3, 50, -2, 9), (1, 64, 2, 200, 0, 191),
                                           (1, 100, 2, 77, 5, 104), (2, 70, 2, 45, -10, 80)])
def test_stereo_sad_backward(N, C, H, W, mn, mx):
    from bplayers_amd import pytorch_cuda_stereo_sad_op as op
    rng = np.random.default_rng(C)
    f0 = rng.normal(size=(N, C, H, W)).astype(np.float32)
    f1 = rng.normal(size=(N, C, H, W)).astype(np.float32)
    g = rng.normal(size=(N, H, W, mx - mn + 1)).astype(np.float32)
    d0, d1 = op.backward(_t(f0), _t(f1), mn, mx, _t(g))
    e0, e1 = S.stereo_sad_backward(f0, f1, mn, mx, g)
    k0, k1 = S.stereo_sad_backward(f0, f1, mn, mx, g, absolute=True)
    P.conditioned("df0", d0.cpu().numpy(), e0, k0)
    P.conditioned("df1", d1.cpu().numpy(), e1, k1)


@pytest.mark.parametrize("sws,ou,ov,bu,bv", [(24, 0, 0, 0, 0), (24, 3, -2, 1, 1), (8, 0, 0, 0, 0), (10, -5, 5, 0, 1),
                                            (5, 0, 0, 0, 0), (7, 2, -1, 1, 0), (40, 0, 0, 0, 0), (33, 2, -3, 1, 0),
                                            (64, -4, 4, 0, 1)])
def test_flow_sad_forward(sws, ou, ov, bu, bv):
    from bplayers_amd import pytorch_cuda_flow_mp_sad_op as op
    rng = np.random.default_rng(sws)
    f0 = rng.normal(size=(1, 12, 19, 70)).astype(np.float32)
    f1 = rng.normal(size=(1, 12, 19, 70)).astype(np.float32)
    got = [t.cpu().numpy() for t in op.forward(_t(f0), _t(f1), sws, ou, ov, bu, bv)]
    exp = S.flow_sad_forward(f0, f1, sws, ou, ov, bu, bv, want_volume=True)
    check_flow_forward(got, exp, f0.shape[1], sws, bu, bv)


def check_flow_forward(got, exp, C, sws, bu=0, bv=0):
    """Volumes elementwise; argmins (first index of the minimum) identical except where the
    GPU's label attains a value within the fp32 bound of a sum of C non-negative terms,
    2 C 2^-24 S, of the oracle's minimum (a rounding tie)."""
    P.elementwise("cv_u", got[0], exp[0])
    P.elementwise("cv_v", got[1], exp[1])
    Svol = exp[4]                                # [N,H,W,v_idx,u_idx]
    for name, g, e, axis, blk in (("u_star", got[2], exp[2], 3, bv), ("v_star", got[3], exp[3], 4, bu)):
        gi = np.rint(g).astype(np.int64) - blk * sws
        ei = np.rint(e).astype(np.int64) - blk * sws
        assert gi.min() >= 0 and gi.max() < Svol.shape[axis], f"{name}: label out of range"
        vol = np.moveaxis(Svol, axis, -1)        # [N,H,W,other,label-of-min]
        sg = np.take_along_axis(vol, gi[..., None], axis=-1)[..., 0]
        se = np.take_along_axis(vol, ei[..., None], axis=-1)[..., 0]
        tie = np.abs(sg - se) <= 2 * C * U32 * np.maximum(sg, se)
        bad = (gi != ei) & ~tie
        excl = float(((gi != ei) & tie).mean())
        print(f"    {name}: {int(bad.sum())} mismatches outside fp32 ties; {excl:.2e} rounding ties", flush=True)
        assert bad.sum() == 0, f"{name}: {int(bad.sum())} argmins differ outside rounding ties"
        assert excl <= P.ARGMIN_EXCLUDED_CAP


@pytest.mark.parametrize("N,C,H,W,sws,stars", [
    (1, 6, 30, 40, 24, "fwd"),
    (1, 37, 45, 77, 8, "fwd"),      # ragged channel groups, partial tiles
    (2, 16, 19, 70, 5, "fwd"),      # odd window: the last label has no gradient
    (1, 4, 20, 33, 31, "fwd"),      # K = 32, the largest masked window
    (1, 3, 18, 25, 34, "fwd"),      # K = 35: searched landings
    (1, 5, 17, 29, 6, "offset"),    # block-offset argmins land outside the window
    (2, 3, 13, 21, 5, "random"),    # arbitrary argmins, negatives included
    (1, 2, 12, 40, 33, "random"),
    # sources converging on every K-th row (u volume) and column (v volume): ~2h extra landings
    # per label on those targets, past the block's pool of extra landings
    (1, 4, 30, 40, 24, "converge"),
    (2, 3, 21, 37, 8, "converge"),
])
def test_flow_sad_backward(N, C, H, W, sws, stars):
    from bplayers_amd import pytorch_cuda_flow_mp_sad_op as op
    rng = np.random.default_rng(7 + sws)
    K = sws + 1
    f0 = rng.normal(size=(N, C, H, W)).astype(np.float32)
    f1 = rng.normal(size=(N, C, H, W)).astype(np.float32)
    if stars == "fwd":
        _, _, us, vs = S.flow_sad_forward(f0, f1, sws)
    elif stars == "offset":
        _, _, us, vs = S.flow_sad_forward(f0, f1, sws, 2, -1, 1, 1)
    elif stars == "converge":
        hs = sws // 2
        ys = np.arange(H)[:, None, None]
        xs = np.arange(W)[None, :, None]
        # label s of the other displacement e = s - hs lands row ys + e (u volume) or column
        # xs + e (v volume) on the nearest target line hs + K * m
        us = np.broadcast_to(np.clip(hs + ((ys // K) * K + hs - ys), 0, 2 * hs), (N, H, W, K)).astype(np.float64)
        vs = np.broadcast_to(np.clip(hs + ((xs // K) * K + hs - xs), 0, 2 * hs), (N, H, W, K)).astype(np.float64)
    else:
        us = rng.integers(-3, 2 * K + 2, size=(N, H, W, K)).astype(np.float64)
        vs = rng.integers(-3, 2 * K + 2, size=(N, H, W, K)).astype(np.float64)
    us = us.astype(np.float32)
    vs = vs.astype(np.float32)
    gu = rng.normal(size=(N, H, W, K)).astype(np.float32)
    gv = rng.normal(size=(N, H, W, K)).astype(np.float32)
    d0, d1 = op.backward(_t(f0), _t(f1), sws, _t(gu), _t(gv), _t(us), _t(vs))
    e0, e1 = S.flow_sad_backward(f0, f1, sws, gu, gv, us, vs)
    k0, k1 = S.flow_sad_backward(f0, f1, sws, gu, gv, us, vs, absolute=True)
    P.conditioned("df0", d0.cpu().numpy(), e0, k0)
    P.conditioned("df1", d1.cpu().numpy(), e1, k1)
    np.testing.assert_allclose(d1.cpu().numpy(), e1, rtol=1e-4, atol=1e-4)


def test_flow_sad_backward_deterministic():
    """In-window argmins: df1 is a fixed-order gather, bitwise repeatable."""
    from bplayers_amd import pytorch_cuda_flow_mp_sad_op as op
    rng = np.random.default_rng(5)
    sws = 24
    f0 = rng.normal(size=(1, 32, 40, 70)).astype(np.float32)
    f1 = rng.normal(size=(1, 32, 40, 70)).astype(np.float32)
    cu, cv, us, vs = op.forward(_t(f0), _t(f1), sws, 0, 0, 0, 0)
    gu = torch.randn_like(cu)
    gv = torch.randn_like(cv)
    a = op.backward(_t(f0), _t(f1), sws, gu, gv, us, vs)
    b = op.backward(_t(f0), _t(f1), sws, gu, gv, us, vs)
    assert torch.equal(a[0], b[0]) and torch.equal(a[1], b[1])


def test_flow_sad_function_sws216_matches_oracle():
    """FlowMpSadFunction at sws = 216 (one 217 x 217 window here; the reference tiles it into
    four 109-windows whose merge gives the same minima and first-index argmins, pinned on the
    CPU in test_oracle_sad.py) against the FP64 oracle's 217-window: volumes, argmins and the
    feature gradients of a weighted sum of both volumes."""
    from bplayers_amd.ops.flow_mp_sad.flow_mp_sad import FlowMpSadFunction
    rng = np.random.default_rng(9)
    a0 = rng.normal(size=(1, 3, 12, 14)).astype(np.float32)
    a1 = rng.normal(size=(1, 3, 12, 14)).astype(np.float32)
    f0, f1 = _t(a0).requires_grad_(), _t(a1).requires_grad_()
    cu, cv, us, vs = FlowMpSadFunction.apply(f0, f1, 216)
    exp = S.flow_sad_forward(a0, a1, 216, want_volume=True)
    check_flow_forward([x.detach().cpu().numpy() for x in (cu, cv, us, vs)], exp, 3, 216)
    gu = rng.normal(size=cu.shape).astype(np.float32)
    gv = rng.normal(size=cv.shape).astype(np.float32)
    ((cu * _t(gu)).sum() + (cv * _t(gv)).sum()).backward()
    us64, vs64 = us.cpu().numpy(), vs.cpu().numpy()   # the routing the GPU replays
    e0, e1 = S.flow_sad_backward(a0, a1, 216, gu, gv, us64, vs64)
    k0, k1 = S.flow_sad_backward(a0, a1, 216, gu, gv, us64, vs64, absolute=True)
    P.conditioned("df0", f0.grad.cpu().numpy(), e0, k0)
    P.conditioned("df1", f1.grad.cpu().numpy(), e1, k1)


def test_stereo_matching_sad_function_gradients_match_oracle():
    """StereoMatchingSadFunction (ops/sad/stereo_sad.py:5-19): cost volume and the feature
    gradients of sum(G * cost) against the FP64 oracle."""
    from bplayers_amd.ops.sad.stereo_sad import StereoMatchingSadFunction
    rng = np.random.default_rng(21)
    N, C, H, W, mn, mx = 1, 16, 5, 90, -2, 61
    a0 = rng.normal(size=(N, C, H, W)).astype(np.float32)
    a1 = rng.normal(size=(N, C, H, W)).astype(np.float32)
    G = rng.normal(size=(N, H, W, mx - mn + 1)).astype(np.float32)
    f0, f1 = _t(a0).requires_grad_(), _t(a1).requires_grad_()
    cost = StereoMatchingSadFunction.apply(f0, f1, mn, mx)
    (cost * _t(G)).sum().backward()
    P.elementwise("SAD volume", cost.detach().cpu().numpy(), S.stereo_sad_forward(a0, a1, mn, mx))
    e0, e1 = S.stereo_sad_backward(a0, a1, mn, mx, G)
    k0, k1 = S.stereo_sad_backward(a0, a1, mn, mx, G, absolute=True)
    P.conditioned("df0", f0.grad.cpu().numpy(), e0, k0)
    P.conditioned("df1", f1.grad.cpu().numpy(), e1, k1)


def test_stereo_matching_layer_gradients_match_oracle():
    """StereoMatchingSad (fused SAD + softmax(-SAD / T)) against the FP64 composition: the
    probabilities, and the gradients of sum(G * prob) to f0, f1 and T (softmax Jacobian in
    FP64, then the oracle's SAD backward)."""
    from bplayers_amd.matching import StereoMatchingSad
    rng = np.random.default_rng(3)
    N, C, H, W, mn, mx, T = 1, 8, 4, 70, 0, 47, 2.5
    a0 = rng.normal(size=(N, C, H, W)).astype(np.float32)
    a1 = rng.normal(size=(N, C, H, W)).astype(np.float32)
    G = rng.normal(size=(N, H, W, mx - mn + 1))
    layer = StereoMatchingSad(mn, mx, init_temp=T, device="cuda:0")
    f0, f1 = _t(a0).requires_grad_(), _t(a1).requires_grad_()
    prob = layer(f0, f1)
    (prob * _t(G)).sum().backward()
    sad = S.stereo_sad_forward(a0, a1, mn, mx)
    z = -sad / T
    p = np.exp(z - z.max(-1, keepdims=True))
    p /= p.sum(-1, keepdims=True)
    P.elementwise("probabilities", prob.detach().cpu().numpy(), p)
    gz = p * (G - (G * p).sum(-1, keepdims=True))          # dL/dz
    kz = p * (np.abs(G) + (np.abs(G) * p).sum(-1, keepdims=True))
    gs, ks = -gz / T, kz / T                                  # dL/dSAD and its magnitude
    e0, e1 = S.stereo_sad_backward(a0, a1, mn, mx, gs)
    k0, k1 = S.stereo_sad_backward(a0, a1, mn, mx, ks, absolute=True)
    P.conditioned("df0", f0.grad.cpu().numpy(), e0, k0)
    P.conditioned("df1", f1.grad.cpu().numpy(), e1, k1)
    gT = (gz * sad).sum() / T ** 2
    kT = (kz * sad).sum() / T ** 2
    P.conditioned("dT", layer._softmax.T.grad.cpu().numpy().reshape(()), np.array(gT), np.array(kT))


@pytest.mark.parametrize("N,C,H,W,mn,mx", [(1, 16, 5, 100, 0, 63), (2, 24, 3, 77, -3, 96), (1, 64, 2, 200, 0, 191),
                                          (1, 8, 4, 130, 2, 257), (1, 3, 2, 40, 0, 9)])
def test_stereo_sad_softmax_fused_equals_two_kernels(N, C, H, W, mn, mx):
    """The fused SAD + softmax kernel equals the SAD kernel followed by the label softmax bit
    for bit (same per-thread SAD arithmetic, same softmax trees), for both signs."""
    from bplayers_amd import pytorch_cuda_stereo_sad_op as op
    from bplayers_amd import _native as nat
    rng = np.random.default_rng(C + W)
    f0 = _t(rng.normal(size=(N, C, H, W)))
    f1 = _t(rng.normal(size=(N, C, H, W)))
    D = mx - mn + 1
    sad = op.forward(f0, f1, mn, mx)
    for scale, sign in ((-1.0 / 3.0, 1.0), (-0.25, -1.0)):
        got = op.forward_softmax(f0, f1, mn, mx, scale, sign)
        exp = torch.empty_like(sad)
        nat.call("bpl_softmax_labels", nat.ptr(sad), nat.ptr(exp), N * H * W, D, scale, sign, nat.stream_of(sad))
        assert torch.equal(got, exp)
        # and the FP64 oracle
        s64 = S.stereo_sad_forward(f0.cpu().numpy(), f1.cpu().numpy(), mn, mx) * scale
        p64 = np.exp(s64 - s64.max(axis=3, keepdims=True))
        p64 = sign * p64 / p64.sum(axis=3, keepdims=True)
        # condition magnitudes of a softmax output under relative errors of its SAD inputs
        q = np.abs(p64)
        z = np.abs(s64)
        P.conditioned(f"fused SAD softmax, scale {scale:.3g}", got.cpu().numpy(), p64,
                      q * (z + (q * z).sum(axis=3, keepdims=True)))
