"""CPU tests of the semantic front-end (PWNet / EdgeNet and their U-Net): padding, shapes,
strict loading of the reference's semantic checkpoints, and properties of the architecture.
The networks' outputs themselves are parity unpinned: the reference holds no front-end outputs
(its only semantic artefact needs the absent ESPNet)."""
import os
import types

import pytest
import torch
import torch.nn.functional as F

from bplayers_amd.corenet import StereoUnaryUnetDyn, crop_padding, pad_to_multiple
from bplayers_amd.networks import EdgeNet, PWNet, strip_duplicate_unet

REF_SEM = "/root/reference/data/params/semantic"


def _args(**kw):
    a = dict(num_labels=20, num_bp_layers=1, activation='leakyrelu', with_bn=True, with_upconv=False,
             with_output_bn=False, multi_level_output=False, pad=(0, 0), pairwise_type='pixel', with_edges=False)
    a.update(kw)
    return types.SimpleNamespace(**a)


@pytest.mark.parametrize("H,W,extra", [(37, 50, (0, 0)), (40, 48, (0, 0)), (33, 41, (2, 3))])
def test_padding_split_and_round_trip(H, W, extra):
    x = torch.randn(1, 3, H, W)
    p, (l, r, t, b) = pad_to_multiple(x, 8, extra)
    assert p.shape[2] % 8 == (2 * extra[0]) % 8 and p.shape[3] % 8 == (2 * extra[1]) % 8
    dh, dw = p.shape[2] - H - 2 * extra[0], p.shape[3] - W - 2 * extra[1]
    assert (H + dh) % 8 == 0 and (W + dw) % 8 == 0 and 0 <= dh < 8 and 0 <= dw < 8
    # the odd row / column goes to the top / left
    assert t - extra[0] == (dh + 1) // 2 and b - extra[0] == dh // 2
    assert l - extra[1] == (dw + 1) // 2 and r - extra[1] == dw // 2
    assert torch.equal(p, F.pad(x, (l, r, t, b), mode='reflect'))
    assert torch.equal(crop_padding(p, (l, r, t, b)), x)


@pytest.mark.parametrize("upconv,multi,outbn", [(False, False, False), (True, False, True), (False, True, False),
                                                (True, True, True)])
def test_unet_shapes(upconv, multi, outbn):
    torch.manual_seed(0)
    net = StereoUnaryUnetDyn(multi, 'leakyrelu', True, upconv, outbn)
    out = net(torch.randn(2, 3, 24, 40))
    assert len(out) == (3 if multi else 1)
    assert out[0].shape == (2, 32, 24, 40)
    if multi:
        assert out[1].shape == (2, 32, 12, 20) and out[2].shape == (2, 32, 6, 10)
    assert hasattr(net, 'bn12') and (net.bn12 is not None) == outbn


def test_unet_is_invariant_to_positive_input_scale():
    """Every convolution before a per-channel GroupNorm has no bias, so scaling the image by
    s > 0 leaves the U-Net's features unchanged (up to the norm's epsilon)."""
    torch.manual_seed(1)
    net = StereoUnaryUnetDyn(False, 'leakyrelu', True, False, False).double()
    x = torch.randn(1, 3, 16, 24, dtype=torch.float64)
    a, b = net(x)[0], net(3.5 * x)[0]
    assert torch.allclose(a, b, rtol=1e-4, atol=1e-4)
    assert not torch.allclose(a, net(x.flip(-1))[0].flip(-1), atol=1e-3)  # not trivially symmetric


def test_front_end_shapes_and_sign():
    torch.manual_seed(2)
    x = torch.randn(1, 3, 37, 50)
    pw = PWNet("cpu", _args())(x)
    ed = EdgeNet("cpu", _args(pairwise_type='global', with_edges=True))(x)
    assert len(pw) == 1 and pw[0].shape == (1, 2 * 20 * 20, 37, 50) and bool((pw[0] >= 0).all())
    assert len(ed) == 1 and ed[0].shape == (1, 2, 37, 50) and bool((ed[0] >= 0).all())


def test_duplicate_unet_entries_are_checked():
    sd = {"_pw_net.net.a": torch.ones(2), "_pw_net.pad_unpad.net.a": torch.ones(2), "x": torch.zeros(1)}
    assert set(strip_duplicate_unet(sd, "_pw_net.")) == {"_pw_net.net.a", "x"}
    sd["_pw_net.pad_unpad.net.a"] = torch.zeros(2)
    with pytest.raises(ValueError):
        strip_duplicate_unet(sd, "_pw_net.")


@pytest.mark.skipif(not os.path.isdir(REF_SEM), reason="reference checkpoints absent")
@pytest.mark.parametrize("ckpt,kind", [("pixel_model.cpt", "pixel"), ("global_model.cpt", "global")])
def test_reference_semantic_checkpoints_load_strictly(ckpt, kind):
    """The shipped checkpoints fix the architecture: every front-end tensor has a place of the
    same shape here and nothing is left over (load_state_dict strict)."""
    sd = torch.load(os.path.join(REF_SEM, ckpt), map_location="cpu", weights_only=True)
    pre = "_pw_net." if kind == "pixel" else "_edge_net."
    sub = {k[len(pre):]: v for k, v in strip_duplicate_unet(sd, pre).items() if k.startswith(pre)}
    net = (PWNet if kind == "pixel" else EdgeNet)("cpu", _args(pairwise_type=kind, with_edges=kind == "global"))
    net.load_state_dict(sub)  # strict
    out = net(torch.randn(1, 3, 29, 45))[0]
    assert out.shape[2:] == (29, 45) and torch.isfinite(out).all() and bool((out >= 0).all())
    if kind == "global":
        assert tuple(sd["_pw_mat"].shape) == (1, 2, 20, 20)


@pytest.mark.skipif(not os.path.isdir(REF_SEM), reason="reference checkpoints absent")
@pytest.mark.parametrize("ckpt,kind", [("pixel_model.cpt", "pixel"), ("global_model.cpt", "global")])
def test_semantic_net_loads_reference_checkpoint(ckpt, kind):
    from bplayers_amd.semantic_segmentation import SemanticNet
    net = SemanticNet("cpu", _args(pairwise_type=kind, with_edges=kind == "global"))
    skipped = net.load_parameters(os.path.join(REF_SEM, ckpt))
    sd = torch.load(os.path.join(REF_SEM, ckpt), map_location="cpu", weights_only=True)
    assert skipped == sum(k.startswith("_esp_net.") for k in sd) > 0
    assert torch.equal(net._softmax_input.T.detach(), sd["_softmax_input.T"])
    if kind == "global":
        assert torch.equal(net._pw_mat.detach(), sd["_pw_mat"]) and net._edge_net is not None
    else:
        assert torch.equal(net._pw_net.conv_out_lvl0.conv.bias.detach(), sd["_pw_net.conv_out_lvl0.conv.bias"])
