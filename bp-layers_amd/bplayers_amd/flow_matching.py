"""SAD optical-flow matching layer (reference flow_matching.py:9-52): the flow SAD over a
(sws+1)^2 window, min-projected to one volume per flow component (FlowMpSadFunction), negated
and turned into probabilities softmax(-SAD / T) over each volume's labels with one learnable
temperature T (corenet.py:97-104); the two volumes are stacked as [N, 2, H, W, K] for the
5-D BP path.  The parameter name ``_softmax.T`` is the reference's, so its flow matching
checkpoints load (formats.load_checkpoint)."""
import torch
import torch.nn as nn

from .corenet import TemperatureSoftmax
from .ops.flow_mp_sad.flow_mp_sad import FlowMpSadFunction


class FlowMatchingSad(nn.Module):
    def __init__(self, sws, init_temp=1.0, device=None, lvl=0):
        nn.Module.__init__(self)
        self.sws = int(sws)
        self.level = lvl
        self._softmax = TemperatureSoftmax(dim=3, init_temp=init_temp, device=device)

    def compute_score_volume(self, f0, f1):
        """(-cv_u, -cv_v), each [N, H, W, sws+1] (flow_matching.py:50-52)."""
        cv_u, cv_v, _, _ = FlowMpSadFunction.apply(f0.contiguous(), f1.contiguous(), self.sws)
        return -cv_u, -cv_v

    def forward(self, f0, f1):
        score_u, score_v = self.compute_score_volume(f0, f1)
        prob_u = self._softmax(score_u)
        prob_v = self._softmax(score_v)
        return torch.stack((prob_u, prob_v), dim=1).contiguous()

    @staticmethod
    def argmin_to_disp(argmin, min_disp):
        return argmin + min_disp
