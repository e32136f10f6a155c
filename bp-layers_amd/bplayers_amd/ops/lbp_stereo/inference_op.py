"""Inference over beliefs (reference ops/lbp_stereo/inference_op.py:8-97)."""
import torch
import torch.nn as nn


class Inference(nn.Module):
    """modes: wta (argmax), expectation, norm, raw, sub-exp (expectation in a +-3 window
    around the argmax)."""

    def __init__(self, device, mode='wta', mode_passing='min-sum'):
        super(Inference, self).__init__()
        self.device = device
        if mode not in ('wta', 'expectation', 'norm', 'raw', 'sub-exp'):
            raise ValueError("Unknown inference mode " + mode)
        self.mode = mode
        self.mode_passing = mode_passing

    def forward(self, beliefs):
        if self.mode == "wta":
            # the reference's float() branch is overwritten by a second argmax: int64 labels
            res = torch.argmax(beliefs, dim=3, keepdim=True)
        elif self.mode == "expectation":
            beliefs_normal = beliefs / beliefs.sum(dim=3, keepdim=True)
            labels = torch.arange(beliefs.shape[3], device=beliefs.device, dtype=torch.float32).view(1, 1, 1, -1)
            res = (beliefs_normal * labels).sum(dim=3, keepdim=True)
        elif self.mode == "norm":
            res = beliefs / beliefs.sum(dim=3, keepdim=True)
        elif self.mode == "raw":
            res = beliefs
        else:
            res = self.compute_sub_expectation(beliefs)
        return res

    @staticmethod
    def compute_sub_expectation(beliefs, support=3):
        """Expectation of the label over [argmax-support, argmax+support] (clamped), with the
        beliefs renormalised inside the window (inference_op.py:52-97)."""
        N, H, W, K = beliefs.shape
        disp = beliefs.argmax(dim=-1, keepdim=True)
        offs = torch.arange(-support, support + 1, device=beliefs.device)
        win = torch.clamp(disp + offs, 0, K - 1)
        bw = torch.gather(beliefs, -1, win)
        bw = bw / bw.sum(dim=-1, keepdim=True)
        return torch.sum(bw * win, dim=-1, keepdim=True)
