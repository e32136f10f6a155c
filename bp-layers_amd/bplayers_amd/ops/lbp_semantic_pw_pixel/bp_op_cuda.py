"""BP layer for semantic segmentation with a per-pixel pairwise matrix
(reference ops/lbp_semantic_pw_pixel/bp_op_cuda.py:12-131)."""
import torch
import torch.nn as nn

from .._pairwise import smoothness_prob, truncated_linear
from ..lbp_stereo.bp_op_cuda import BP as _StereoBP
from .inference_op import Inference
from .message_passing_op_pw_pixel import MessagePassing


class BP(nn.Module):
    construct_pw_energy_weights = staticmethod(truncated_linear)
    construct_pw_prob_weights = staticmethod(smoothness_prob)
    adjust_input_weights = _StereoBP.adjust_input_weights

    def __init__(self, device, args=None, max_iter=1, num_labels=None, delta=None, mode_inference='expectation',
                 mode_message_passing='min-sum', layer_idx=0, grad_mode='reference'):
        super(BP, self).__init__()
        self.device = device
        self.args = args
        self.max_iter = max_iter
        self.layer_idx = layer_idx
        self.delta = delta if delta is not None else num_labels
        if mode_inference not in ('wta', 'expectation', 'norm', 'raw') or mode_message_passing != 'min-sum':
            raise ValueError("Unknown inference/message passing mode " + mode_inference + " " + mode_message_passing)
        self.message_passing = MessagePassing(self.device, self.max_iter, num_labels, self.delta,
                                              mode_message_passing, grad_mode=grad_mode)
        self.inference = Inference(self.device, mode_inference, mode_passing=mode_message_passing)

    def forward(self, prob_vol, edge_weights, jump):
        N, H, W, K = prob_vol.shape
        messages = torch.zeros((N, 4, H, W, K), requires_grad=False, device=prob_vol.device, dtype=torch.float)
        beliefs = self.message_passing.forward(prob_vol, edge_weights, messages, jump)
        result = self.inference.forward(beliefs)
        return result.permute(0, 3, 1, 2), beliefs

    def project_jumpcosts(self):
        self.message_passing.projectL1L2()

    @property
    def L1(self):
        return self.message_passing.L1

    @L1.setter
    def L1(self, value):
        self.message_passing.L1.data = torch.tensor(value, device=self.device, dtype=torch.float)

    @property
    def L2(self):
        return self.message_passing.L2

    @L2.setter
    def L2(self, value):
        self.message_passing.L2.data = torch.tensor(value, device=self.device, dtype=torch.float)

    @property
    def rescaleT(self):
        return self.message_passing.rescaleT

    @rescaleT.setter
    def rescaleT(self, value):
        self.message_passing.rescaleT = value
