"""Semantic front-end networks (reference networks.py:9-67, 168-213): ``PWNet`` predicts the
per-pixel C x C pairwise costs consumed by the per-pixel semantic LBP op, ``EdgeNet`` the
per-pixel edge weights of the global-matrix variant.  Both are the U-Net of corenet plus one
reflect-padded 3x3 output convolution per level, with |.| applied to the output.

Module and parameter names follow the reference's state_dict keys (``net.*``,
``conv_out_lvl<i>.conv.*``), so the shipped semantic checkpoints' ``_pw_net.*`` /
``_edge_net.*`` entries load strictly.  The reference registers its U-Net twice (as ``net`` and
as ``pad_unpad.net``); its checkpoints therefore carry every U-Net tensor twice, and loading
here checks that the two copies are equal and keeps one.  Checkpoints are read with
``torch.load(..., weights_only=True)``."""
import os.path as osp

import torch
import torch.nn as nn

from .corenet import PaddedConv2d, StereoUnaryUnetDyn, crop_padding, pad_to_multiple

_DUP = "pad_unpad."


def strip_duplicate_unet(state_dict, prefix=""):
    """Drop the ``<prefix>pad_unpad.net.*`` copies of ``<prefix>net.*`` after checking that they
    hold the same values."""
    out = {}
    for k, v in state_dict.items():
        if k.startswith(prefix + _DUP):
            twin = prefix + k[len(prefix) + len(_DUP):]
            if twin not in state_dict or not torch.equal(state_dict[twin], v):
                raise ValueError(f"checkpoint entry {k} has no equal twin {twin}")
            continue
        out[k] = v
    return out


class SubNetwork(nn.Module):
    """A U-Net plus per-level output convolutions, run on a reflect-padded input (padded to the
    U-Net's divisor plus ``args.pad``) and cropped back (networks.py:9-67)."""

    def __init__(self, args, device):
        super().__init__()
        self.args = args
        self.device = device
        self.net = None
        self._pad_input = tuple(getattr(args, 'pad', (0, 0)))

    def features(self, ipt):
        padded, lrtb = pad_to_multiple(ipt, self.net.divisor, self._pad_input)
        return [crop_padding(o, lrtb) for o in self.net(padded)]

    def forward(self, ipt):
        return self.features(ipt)

    def freeze_parameters(self):
        self.eval()
        for p in self.parameters():
            p.requires_grad = False

    def parameter_list(self, requires_grad=None):
        return [p for p in self.parameters() if requires_grad is None or p.requires_grad == requires_grad]

    def hook_adjust_checkpoint(self, checkpoint):
        return checkpoint

    def load_parameters(self, path, device=None):
        if path is None:
            return
        if not osp.exists(path):
            raise FileNotFoundError("Specified unary checkpoint file does not exist!" + path)
        sd = torch.load(path, map_location=device if device is not None else 'cpu', weights_only=True)
        sd = strip_duplicate_unet(self.hook_adjust_checkpoint(sd))
        self.load_state_dict(sd)

    @property
    def divisor(self):
        return self.net.divisor

    @property
    def pad_input(self):
        return self._pad_input


def _unet(args):
    return StereoUnaryUnetDyn(getattr(args, 'multi_level_output', False), getattr(args, 'activation', 'leakyrelu'),
                              getattr(args, 'with_bn', True), getattr(args, 'with_upconv', False),
                              getattr(args, 'with_output_bn', False))


class _OutputHeads(SubNetwork):
    def __init__(self, device, args, out_channels):
        super().__init__(args, device)
        self.net = _unet(args)
        for lvl, ic in enumerate(self.net.out_channels):
            self.add_module(f"conv_out_lvl{lvl}", PaddedConv2d(ic, out_channels, 3, bias=True))
        self.to(device)

    def forward(self, ipt):
        return [torch.abs(getattr(self, f"conv_out_lvl{lvl}")(f)) for lvl, f in enumerate(self.features(ipt))]


class EdgeNet(_OutputHeads):
    """Edge weights |conv(features)|, 2 channels (x, y) per BP layer (networks.py:168-187)."""

    def __init__(self, device, args):
        super().__init__(device, args, args.num_bp_layers * 2)


class PWNet(_OutputHeads):
    """Per-pixel pairwise costs |conv(features)|, 2 * C * C channels per BP layer: for BP layer 0,
    channel (d, l, c) = pairwise cost of labels (l, c) in direction d (0 horizontal, 1 vertical)
    (networks.py:190-213)."""

    def __init__(self, device, args):
        super().__init__(device, args, args.num_bp_layers * 2 * args.num_labels * args.num_labels)
