"""GPU parity of the SAD ops against the FP64 oracle."""
import numpy as np
import pytest
import torch

from oracle import sad as S

pytestmark = pytest.mark.gpu

SAD_RTOL = 1e-5  # fp32 sum over channels vs float64


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
    np.testing.assert_allclose(got, exp, rtol=SAD_RTOL, atol=1e-5)


def test_stereo_sad_forward_fractional_step():
    from bplayers_amd import pytorch_cuda_stereo_sad_op as op
    rng = np.random.default_rng(1)
    f0 = rng.normal(size=(1, 4, 3, 20)).astype(np.float32)
    f1 = rng.normal(size=(1, 4, 3, 20)).astype(np.float32)
    got = op.forward(_t(f0), _t(f1), 0, 9, 0.5).cpu().numpy()
    np.testing.assert_allclose(got, S.stereo_sad_forward(f0, f1, 0, 9, 0.5), rtol=SAD_RTOL, atol=1e-5)


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
    np.testing.assert_allclose(d0.cpu().numpy(), e0, rtol=1e-4, atol=1e-4)
    np.testing.assert_allclose(d1.cpu().numpy(), e1, rtol=1e-4, atol=1e-4)


@pytest.mark.parametrize("sws,ou,ov,bu,bv", [(24, 0, 0, 0, 0), (24, 3, -2, 1, 1), (8, 0, 0, 0, 0), (10, -5, 5, 0, 1),
                                            (5, 0, 0, 0, 0), (7, 2, -1, 1, 0), (40, 0, 0, 0, 0), (33, 2, -3, 1, 0),
                                            (64, -4, 4, 0, 1)])
def test_flow_sad_forward(sws, ou, ov, bu, bv):
    from bplayers_amd import pytorch_cuda_flow_mp_sad_op as op
    rng = np.random.default_rng(sws)
    f0 = rng.normal(size=(1, 12, 19, 70)).astype(np.float32)
    f1 = rng.normal(size=(1, 12, 19, 70)).astype(np.float32)
    got = [t.cpu().numpy() for t in op.forward(_t(f0), _t(f1), sws, ou, ov, bu, bv)]
    exp = S.flow_sad_forward(f0, f1, sws, ou, ov, bu, bv)
    np.testing.assert_allclose(got[0], exp[0], rtol=SAD_RTOL, atol=1e-5)
    np.testing.assert_allclose(got[1], exp[1], rtol=SAD_RTOL, atol=1e-5)
    # argmins: identical where the minimum is not a near-tie in the oracle
    Sfull = None
    for g, e in ((got[2], exp[2]), (got[3], exp[3])):
        mism = g != e
        assert mism.mean() < 1e-3, f"argmin mismatches {mism.sum()}"


@pytest.mark.parametrize("N,C,H,W,sws,stars", [
    (1, 6, 30, 40, 24, "fwd"),
    (1, 37, 45, 77, 8, "fwd"),      # ragged channel groups, partial tiles
    (2, 16, 19, 70, 5, "fwd"),      # odd window: the last label has no gradient
    (1, 4, 20, 33, 31, "fwd"),      # K = 32, the largest masked window
    (1, 3, 18, 25, 34, "fwd"),      # K = 35: searched landings
    (1, 5, 17, 29, 6, "offset"),    # block-offset argmins land outside the window
    (2, 3, 13, 21, 5, "random"),    # arbitrary argmins, negatives included
    (1, 2, 12, 40, 33, "random"),
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
    else:
        us = rng.integers(-3, 2 * K + 2, size=(N, H, W, K)).astype(np.float64)
        vs = rng.integers(-3, 2 * K + 2, size=(N, H, W, K)).astype(np.float64)
    us = us.astype(np.float32)
    vs = vs.astype(np.float32)
    gu = rng.normal(size=(N, H, W, K)).astype(np.float32)
    gv = rng.normal(size=(N, H, W, K)).astype(np.float32)
    d0, d1 = op.backward(_t(f0), _t(f1), sws, _t(gu), _t(gv), _t(us), _t(vs))
    e0, e1 = S.flow_sad_backward(f0, f1, sws, gu, gv, us, vs)
    np.testing.assert_allclose(d0.cpu().numpy(), e0, rtol=1e-4, atol=1e-4)
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


def test_flow_sad_function_sws216_matches_tiled_merge():
    """FlowMpSadFunction computes sws=216 as one window; the reference tiles it into four
    109-windows merged with '<'.  The two give the same volumes and argmins."""
    from bplayers_amd.ops.flow_mp_sad.flow_mp_sad import FlowMpSadFunction
    from bplayers_amd import pytorch_cuda_flow_mp_sad_op as op
    rng = np.random.default_rng(9)
    f0 = _t(rng.normal(size=(1, 3, 12, 14)))
    f1 = _t(rng.normal(size=(1, 3, 12, 14)))
    cu, cv, us, vs = FlowMpSadFunction.apply(f0, f1, 216)
    s = 108
    t = {}
    for key, (ou, ov, bu, bv) in {"00": (-s // 2, -s // 2, 0, 0), "01": (-s // 2, s // 2, 0, 1),
                                   "10": (s // 2, -s // 2, 1, 0), "11": (s // 2, s // 2, 1, 1)}.items():
        t[key] = op.forward(f0, f1, s, ou, ov, bu, bv)
    u0 = torch.cat((t["00"][0], t["10"][0][..., 1:]), -1)
    u1 = torch.cat((t["01"][0], t["11"][0][..., 1:]), -1)
    a0 = torch.cat((t["00"][2], t["10"][2][..., 1:]), -1)
    a1 = torch.cat((t["01"][2], t["11"][2][..., 1:]), -1)
    idx = u1 < u0
    u0[idx] = u1[idx]
    a0[idx] = a1[idx]
    assert torch.equal(cu, u0)
    assert torch.equal(us, a0)


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
        np.testing.assert_allclose(got.cpu().numpy(), p64, rtol=1e-4, atol=1e-6)


def test_stereo_matching_layer_fused_gradients_match_composition():
    """StereoMatchingSad (fused forward, softmax Jacobian + SAD backward) against the
    reference's composition SAD op -> negation -> softmax(x / T) under autograd: the
    probabilities and the gradients to f0, f1 and T."""
    from bplayers_amd.matching import StereoMatchingSad
    rng = np.random.default_rng(3)
    N, C, H, W, mn, mx = 1, 8, 4, 70, 0, 47
    a0 = rng.normal(size=(N, C, H, W))
    a1 = rng.normal(size=(N, C, H, W))
    G = _t(rng.normal(size=(N, H, W, mx - mn + 1)))
    layer = StereoMatchingSad(mn, mx, init_temp=2.5, device="cuda:0")
    f0 = _t(a0).requires_grad_()
    f1 = _t(a1).requires_grad_()
    p = layer(f0, f1)
    (p * G).sum().backward()
    T2 = torch.full((1,), 2.5, device="cuda:0", requires_grad=True)
    g0 = _t(a0).requires_grad_()
    g1 = _t(a1).requires_grad_()
    q = torch.softmax(layer.compute_score_volume(g0, g1) / T2, dim=3)
    (q * G).sum().backward()
    torch.testing.assert_close(p, q, rtol=1e-5, atol=1e-7)
    torch.testing.assert_close(f0.grad, g0.grad, rtol=1e-4, atol=1e-5)
    torch.testing.assert_close(f1.grad, g1.grad, rtol=1e-4, atol=1e-5)
    torch.testing.assert_close(layer._softmax.T.grad, T2.grad, rtol=1e-4, atol=1e-5)
