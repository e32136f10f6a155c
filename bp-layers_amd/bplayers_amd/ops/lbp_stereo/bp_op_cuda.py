"""BP layer around the stereo message passing (reference ops/lbp_stereo/bp_op_cuda.py:15-258).

The reference derives BP from its network base class (checkpoint loading through
command-line args); here BP is a plain nn.Module with an optional ``checkpoint`` path that
is loaded with ``torch.load(weights_only=True)`` and the reference's key adjustments."""
import torch
import torch.nn as nn

from .._pairwise import jump_index, jump_value, smoothness_prob, truncated_linear
from .inference_op import Inference
from .message_passing_op_cuda import MessagePassing


class BP(nn.Module):

    # closed-form pairwise tables of the global P1/P2 model and the jump-vector lookup, for
    # callers of the reference's static helpers (bp_op_cuda.py:18-78)
    get_linear_idx = staticmethod(jump_index)
    f_func = staticmethod(jump_value)
    construct_pw_energy_weights = staticmethod(truncated_linear)
    construct_pw_prob_weights = staticmethod(smoothness_prob)

    def __init__(self, device, args=None, max_iter=1, num_labels=None, delta=3, mode_inference='expectation',
                 mode_message_passing='min-sum', layer_idx=0, level=0, grad_mode='reference', checkpoint=None):
        super(BP, self).__init__()
        self.device = device
        self.args = args
        self.max_iter = max_iter
        self.layer_idx = layer_idx
        self.level = level
        self.delta = delta
        if mode_inference not in ('wta', 'expectation', 'norm', 'raw', 'sub-exp') or mode_message_passing != 'min-sum':
            raise ValueError("Unknown inference/message passing mode " + mode_inference + " " + mode_message_passing)
        self.message_passing = MessagePassing(self.device, self.max_iter, num_labels, delta, mode_message_passing,
                                              grad_mode=grad_mode)
        self.inference = Inference(self.device, mode_inference, mode_passing=mode_message_passing)
        if checkpoint is not None:
            self.load_parameters(checkpoint, device)

    def load_parameters(self, path, device):
        ckpt = torch.load(path, map_location=device, weights_only=True)
        self.load_state_dict(self.hook_adjust_checkpoint(ckpt))

    def forward(self, prob_vol, edge_weights, affinities=None, offsets=None):
        if len(prob_vol.shape) == 5:
            N, _, H, W, K = prob_vol.shape
            # u-flow
            messages_u = torch.zeros((N, 4, H, W, K), requires_grad=True, device=prob_vol.device, dtype=torch.float)
            beliefs_u, messages_u = self.message_passing.forward(prob_vol[:, 0].contiguous(), edge_weights,
                                                                 affinities, offsets, messages_u, messages_zero=True)
            result_u = self.inference.forward(beliefs_u)
            # v-flow: starts from the u messages (bp_op_cuda.py:110)
            beliefs_v, messages_v = self.message_passing.forward(prob_vol[:, 1].contiguous(), edge_weights,
                                                                 affinities, offsets, messages_u)
            result_v = self.inference.forward(beliefs_v)
            flow = torch.cat((result_u, result_v), dim=-1).permute(0, 3, 1, 2)
            beliefs = torch.cat((beliefs_u.unsqueeze(1), beliefs_v.unsqueeze(1)), dim=1)
            messages = torch.cat((messages_u.unsqueeze(1), messages_v.unsqueeze(1)), dim=1)
            return flow, beliefs, messages
        N, H, W, K = prob_vol.shape
        # the initial messages are constant zeros: no gradient is wanted for them (the
        # reference marks them requires_grad; nothing can read that gradient)
        messages = torch.zeros((N, 4, H, W, K), requires_grad=False, device=prob_vol.device, dtype=torch.float)
        beliefs, messages = self.message_passing.forward(prob_vol.contiguous(), edge_weights, affinities, offsets,
                                                         messages, messages_zero=True)
        result = self.inference.forward(beliefs)
        disps = result.permute(0, 3, 1, 2)
        return disps, beliefs, messages

    def adjust_input_weights(self, weights, idx):
        """[N,2,H,W] (x, y) edge weights -> [N,4,H,W]; RIGHT/DOWN read the left/upper
        neighbour's weight, zero on the first column/row (bp_op_cuda.py:130-154)."""
        if weights is None:
            return None
        w = torch.zeros((weights.shape[0], 4, weights.shape[2], weights.shape[3]), device=weights.device)
        w[:, 0] = weights[:, 0]
        w[:, 1, :, 1:] = weights[:, 0, :, :-1]
        w[:, 2] = weights[:, 1]
        w[:, 3, 1:, :] = weights[:, 1, :-1, :]
        return w.contiguous()

    def adjust_input_affinities(self, affinities):
        """[N,2,5,H,W] -> ordered (L3 >= L2 >= L1) and shifted [N,4,5,H,W] (bp_op_cuda.py:156-188)."""
        if affinities is None:
            return None
        a = affinities.clone()
        a[:, :, 0] = torch.max(affinities[:, :, 0], affinities[:, :, 2])
        a[:, :, 1] = torch.max(affinities[:, :, 1], affinities[:, :, 3])
        a[:, :, 4] = torch.max(affinities[:, :, 4], torch.max(a[:, :, 0].clone(), a[:, :, 1].clone()))
        s = torch.zeros((a.shape[0], a.shape[1] + 2, a.shape[2], a.shape[3], a.shape[4]), device=a.device)
        s[:, 0] = a[:, 0]
        s[:, 1, :, :, 1:] = a[:, 0, :, :, :-1]
        s[:, 2] = a[:, 1]
        s[:, 3, :, 1:, :] = a[:, 1, :, :-1, :]
        return s.contiguous()

    def adjust_input_offsets(self, offsets):
        """[N,2,H,W] -> [N,4,H,W]; negated and shifted for RIGHT/DOWN (bp_op_cuda.py:190-209)."""
        if offsets is None:
            return None
        s = torch.zeros((offsets.shape[0], offsets.shape[1] + 2, offsets.shape[2], offsets.shape[3]),
                        device=offsets.device)
        s[:, 0] = offsets[:, 0]
        s[:, 1, :, 1:] = -offsets[:, 0, :, :-1]
        s[:, 2] = offsets[:, 1]
        s[:, 3, 1:, :] = -offsets[:, 1, :-1, :]
        return s.contiguous()

    def project_jumpcosts(self):
        self.message_passing.projectL1L2()

    def hook_adjust_checkpoint(self, checkpoint):
        """Older checkpoints: add softmin.T = 1 and rename L1/L2 to _L1/_L2 (bp_op_cuda.py:221-237)."""
        if 'message_passing.softmin.T' not in checkpoint.keys():
            checkpoint['message_passing.softmin.T'] = self.message_passing.softmin.T
        if 'message_passing.L1' in checkpoint.keys() and 'message_passing.L2' in checkpoint.keys():
            checkpoint['message_passing._L1'] = checkpoint.pop('message_passing.L1')
            checkpoint['message_passing._L2'] = checkpoint.pop('message_passing.L2')
        return checkpoint

    @property
    def L1(self):
        return self.message_passing.L1

    @property
    def L2(self):
        return self.message_passing.L2

    def setL1L2(self, value_L1, value_L2):
        self.message_passing.setL1L2(value_L1, value_L2)

    @property
    def rescaleT(self):
        return self.message_passing.rescaleT

    @rescaleT.setter
    def rescaleT(self, value):
        self.message_passing.rescaleT = value
