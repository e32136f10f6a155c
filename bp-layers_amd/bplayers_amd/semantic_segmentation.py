"""Semantic segmentation CRF model around the semantic LBP ops (reference
semantic_segmentation.py:19-145).

The reference computes its unaries with ESPNet (a git submodule with its own pretrained
weights, not available here); ``SemanticNet.forward`` therefore takes the unary logits
[N, C, H, W] from the caller next to the image.  Everything after the unary network is the
reference's pipeline: softmax(logits / T) over labels, then one BP layer with either a global
pairwise matrix ``_pw_mat`` [1,2,C,C] (optionally with per-pixel edge weights from ``EdgeNet``)
or per-pixel pairwise costs from ``PWNet``, reshaped from [1, 2*C*C, H, W] to [2,H,W,C,C].

``load_parameters`` reads the reference's semantic checkpoints (``global_model.cpt``,
``pixel_model.cpt``) with ``torch.load(..., weights_only=True)``: ``_pw_mat``,
``_softmax_input.T`` and the ``_pw_net.*`` / ``_edge_net.*`` networks load strictly; the
``_esp_net.*`` entries (the absent unary network) are skipped and counted."""
import torch
import torch.nn as nn

from .corenet import TemperatureSoftmax
from .networks import EdgeNet, PWNet, strip_duplicate_unet
from .ops.lbp_semantic_pw.bp_op_cuda import BP as BP_PW
from .ops.lbp_semantic_pw_pixel.bp_op_cuda import BP as BP_PW_PIXEL

UNARY_PREFIX = "_esp_net."


class SemanticNet(nn.Module):
    """``args`` fields used (the reference's main_semantic.py options): num_labels,
    pairwise_type ('global' | 'pixel'), with_edges, num_bp_layers, activation, with_bn,
    with_upconv, with_output_bn, multi_level_output, pad.  ``grad_mode`` selects the LBP
    backward ('reference' as the reference, or 'exact')."""

    def __init__(self, device, args, grad_mode='reference'):
        super().__init__()
        self.args = args
        self.device = device
        self.num_labels = C = args.num_labels
        self._softmax_input = TemperatureSoftmax(dim=3, init_temp=1.0, device=device)
        self._edge_net = None
        self._pw_net = None
        self._pw_mat = None
        max_iter = 1
        if args.pairwise_type == "global":
            if getattr(args, 'with_edges', False):
                self._edge_net = EdgeNet(device, args)
            self._crf = [BP_PW(device, args, max_iter, C, C, mode_inference='wta', mode_message_passing='min-sum',
                               layer_idx=0, grad_mode=grad_mode)]
            self._pw_mat = nn.Parameter(torch.zeros((1, 2, C, C), dtype=torch.float, device=device),
                                        requires_grad=True)
        elif args.pairwise_type == "pixel":
            self._crf = [BP_PW_PIXEL(device, args, max_iter, C, C, mode_inference='wta',
                                     mode_message_passing='min-sum', layer_idx=0, grad_mode=grad_mode)]
            self._pw_net = PWNet(device, args)
        else:
            raise ValueError("Unknown pairwise type " + str(args.pairwise_type))

    def load_parameters(self, path):
        """Loads a reference semantic checkpoint; returns the number of skipped unary-network
        entries."""
        sd = torch.load(path, map_location=self.device, weights_only=True)
        skipped = [k for k in sd if k.startswith(UNARY_PREFIX)]
        sd = {k: v for k, v in sd.items() if not k.startswith(UNARY_PREFIX)}
        for pre in ("_pw_net.", "_edge_net."):
            sd = strip_duplicate_unet(sd, pre)
        self.load_state_dict(sd)
        return len(skipped)

    def pairwise_from_net(self, ipt):
        """PWNet output [1, 2*C*C, H, W] -> per-pixel pairwise [2, H, W, C, C] (direction, y, x,
        l, c) (semantic_segmentation.py:132-137; batch 1, as the reference's reshape)."""
        j = self._pw_net(ipt)[0]
        N, _, H, W = j.shape
        if N != 1:
            raise RuntimeError("the per-pixel pairwise reshape supports batch 1 (as the reference)")
        C = self.num_labels
        return j.view((2, C, C, H, W)).permute((0, 3, 4, 1, 2)).contiguous()

    def optimize_crf(self, ipt, prob_vol, weights):
        disps = None
        for idx, crf in enumerate(self._crf):
            prob_vol = prob_vol.contiguous()
            weights_input = crf.adjust_input_weights(weights, idx)
            jump = self.pairwise_from_net(ipt) if self._pw_net is not None else self._pw_mat
            disps, prob_vol = crf.forward(prob_vol, weights_input, jump)
        return disps, prob_vol

    def forward(self, ipt, unary_logits):
        """ipt [N,3,H,W] image, unary_logits [N,C,H,W].  Returns (labels [N,1,H,W] (winner take
        all of the beliefs), beliefs [N,H,W,C], unary-only labels [1,1,H,W] of image 0)."""
        res = self._softmax_input(unary_logits.permute((0, 2, 3, 1)).contiguous())
        unary_only = res[0].max(2)[1].unsqueeze(0).unsqueeze(0)
        N, H, W, C = res.shape
        if self._edge_net is not None:
            weights = self._edge_net(ipt)[0]
        else:
            weights = torch.ones((N, 2, H, W), dtype=torch.float, device=res.device)
        labels, beliefs = self.optimize_crf(ipt, res, weights)
        return labels, beliefs, unary_only
