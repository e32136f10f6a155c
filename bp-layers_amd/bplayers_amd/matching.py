"""SAD stereo matching layer (reference matching.py:7-47): probability volume
softmax(-SAD(f0, f1) / T) over the disparities, with a learnable temperature T
(corenet.py:97-104).  At step 1 and D <= 256 the SAD and the softmax run as one fused kernel;
otherwise as the reference's composition of the SAD op and the temperature softmax."""
import torch
import torch.nn as nn

from .corenet import TemperatureSoftmax
from .ops.sad.stereo_sad import StereoMatchingSadFunction, StereoMatchingSadSoftmaxFunction


class StereoMatchingSad(nn.Module):
    def __init__(self, min_disp, max_disp, step=1.0, init_temp=1.0, device=None):
        nn.Module.__init__(self)
        self.min_disp, self.max_disp, self.step = int(min_disp), int(max_disp), float(step)
        self._softmax = TemperatureSoftmax(dim=3, init_temp=init_temp, device=device)

    def compute_score_volume(self, f0, f1):
        """-SAD (matching.py:45-47)."""
        return -StereoMatchingSadFunction.apply(f0, f1, self.min_disp, self.max_disp, self.step)

    def forward(self, f0, f1):
        D = self.max_disp - self.min_disp + 1
        if self.step == 1.0 and D <= 256:
            return StereoMatchingSadSoftmaxFunction.apply(f0.contiguous(), f1.contiguous(), self.min_disp,
                                                          self.max_disp, self._softmax.T)
        return self._softmax(self.compute_score_volume(f0, f1)).contiguous()

    @staticmethod
    def argmin_to_disp(argmin, min_disp):
        return argmin + min_disp
