"""Stereo min-sum LBP autograd Functions and the MessagePassing layer
(reference ops/lbp_stereo/message_passing_op_cuda.py:1-178).

``LBPMinSumFunction`` keeps the reference's semantics exactly: the message tensor is
updated in place, the saved messages therefore alias across iterations, and the gradient
w.r.t. the input messages is ``in_grad``.  ``LBPMinSumExactFunction`` is the exact
adjoint: each iteration writes fresh messages and returns the true input-message gradient,
so gradients through several iterations match finite differences.
``LBPMinSumExactChainFunction`` runs all iterations of an exact solve as one autograd node,
so its backward can chain the iterations' backward calls (message gradients whose
LEFT/RIGHT slots are zero away from the row ends are neither written nor read there).
``MessagePassing`` selects with ``grad_mode`` ('reference' by default, or 'exact', which uses
the chained Function).
"""
import torch
import torch.nn as nn

from ... import _lbp_op
from ... import _native as nat
from ... import pytorch_cuda_lbp_op_stereo as lbp
from ...corenet import TemperatureSoftmin


def construct_pw_tensor(L1, L2, K):
    """Dense K x K P1/P2 pairwise matrix (message_passing_op_cuda.py:9-20)."""
    pw = torch.ones((K, K), dtype=torch.float) * L2
    for i in range(-1, 2):
        if i == 0:
            pw -= torch.diag(L2 * torch.ones(K - abs(i)), i)
        else:
            pw -= torch.diag((L2 - L1) * torch.ones(K - abs(i)), i)
    return pw.unsqueeze(0).unsqueeze(0)


def pack_affinities(affinities, offset):
    """[N,4,K_in,H,W] affinities + [N,4,H,W] offset -> [N,4,K_in+3,H,W] jump vector
    (message_passing_op_cuda.py:35-56): index 0 the window offset, 1 and -1 the truncation
    (last affinity), the middle 0, then pairs outward; directions 1 and 3 get each pair
    swapped so that their term is the transpose of directions 0 and 2."""
    N, D4, Ki, H, W = affinities.shape
    Kp = Ki + 3
    J = torch.zeros((N, D4, Kp, H, W), dtype=affinities.dtype, device=affinities.device)
    J[:, :, 0] = offset
    J[:, :, 1] = affinities[:, :, -1]
    J[:, :, -1] = affinities[:, :, -1]
    ic = 0
    for i in range(2, Kp // 2):
        J[:, 0::2, i] = affinities[:, 0::2, ic]
        J[:, 0::2, -i] = affinities[:, 0::2, ic + 1]
        J[:, 1::2, i] = affinities[:, 1::2, ic + 1]
        J[:, 1::2, -i] = affinities[:, 1::2, ic]
        ic += 2
    return J.contiguous()


def unpack_affinity_grad(gJ):
    """Adjoint of pack_affinities for the affinities (message_passing_op_cuda.py:75-91);
    returns (grad_affinities, grad_offset)."""
    N, D4, Kp, H, W = gJ.shape
    g = torch.zeros((N, D4, Kp - 3, H, W), dtype=gJ.dtype, device=gJ.device)
    g[:, :, -1] += gJ[:, :, 1]
    g[:, :, -1] += gJ[:, :, -1]
    ic = 0
    for i in range(2, Kp // 2):
        g[:, 0::2, ic] = gJ[:, 0::2, i]
        g[:, 0::2, ic + 1] = gJ[:, 0::2, -i]
        g[:, 1::2, ic + 1] = gJ[:, 1::2, i]
        g[:, 1::2, ic] = gJ[:, 1::2, -i]
        ic += 2
    return g, gJ[:, :, 0]


class LBPMinSumFunction(torch.autograd.Function):
    """One LBP iteration with the reference's semantics (in-place messages)."""

    @staticmethod
    def forward(ctx, cost, edge, messages, delta, affinities, offset):
        jump = pack_affinities(affinities, offset.to(affinities.device))
        messages, messages_argmin, message_scale = lbp.forward_minsum(cost, jump, edge, messages, delta)
        ctx.save_for_backward(cost, jump, edge, messages, messages_argmin, message_scale)
        ctx.delta = int(delta)
        return messages

    @staticmethod
    def backward(ctx, in_grad):
        cost, jump, edge, messages, messages_argmin, message_scale = ctx.saved_tensors
        # grad_messages = in_grad (the reference's semantics) passes through without a copy
        g = in_grad.contiguous()
        grad_cost, grad_jump, grad_edge, grad_message = lbp.backward_minsum(
            cost, jump, edge, g, messages, messages_argmin, message_scale, delta=ctx.delta, grad_messages_out=g)
        grad_aff, grad_offset = unpack_affinity_grad(grad_jump)
        return grad_cost, grad_edge, grad_message, None, grad_aff, grad_offset


class LBPMinSumExactFunction(torch.autograd.Function):
    """One LBP iteration whose backward is the exact adjoint (fresh output messages)."""

    @staticmethod
    def forward(ctx, cost, edge, messages, delta, affinities, offset):
        jump = pack_affinities(affinities, offset.to(affinities.device))
        out = torch.empty_like(messages)
        out, messages_argmin, message_scale = lbp.forward_minsum(cost, jump, edge, messages.contiguous(), delta,
                                                                  out=out)
        ctx.save_for_backward(cost, jump, edge, out, messages_argmin, message_scale)
        ctx.delta = int(delta)
        return out

    @staticmethod
    def backward(ctx, in_grad):
        cost, jump, edge, messages, messages_argmin, message_scale = ctx.saved_tensors
        grad_cost, grad_jump, grad_edge, grad_message = lbp.backward_minsum(
            cost, jump, edge, in_grad.contiguous(), messages, messages_argmin, message_scale, mode="exact",
            delta=ctx.delta)
        grad_aff, grad_offset = unpack_affinity_grad(grad_jump)
        return grad_cost, grad_edge, grad_message, None, grad_aff, grad_offset


class LBPMinSumExactChainFunction(torch.autograd.Function):
    """``max_iter`` (>= 1) iterations with the exact adjoint, as one autograd node: the same
    values as ``max_iter`` applications of LBPMinSumExactFunction, with the iterations'
    backward calls chained (bpl_lbp_bwd's row-end flags)."""

    @staticmethod
    def forward(ctx, cost, edge, messages, delta, affinities, offset, max_iter):
        jump = pack_affinities(affinities, offset.to(affinities.device))
        out, recs = _lbp_op.exact_chain_forward(nat.LBP_STEREO, cost, jump, edge, messages.contiguous(), delta,
                                                int(max_iter))
        ctx.save_for_backward(cost, jump, edge, *[x for r in recs for x in r])
        ctx.delta = int(delta)
        ctx.iters = int(max_iter)
        return out

    @staticmethod
    def backward(ctx, in_grad):
        cost, jump, edge, *flat = ctx.saved_tensors
        recs = [tuple(flat[3 * i:3 * i + 3]) for i in range(ctx.iters)]
        grad_cost, grad_jump, grad_edge, grad_message = _lbp_op.exact_chain_backward(
            nat.LBP_STEREO, cost, jump, edge, in_grad, recs, delta=ctx.delta)
        grad_aff, grad_offset = unpack_affinity_grad(grad_jump)
        return grad_cost, grad_edge, grad_message, None, grad_aff, grad_offset, None


class MessagePassing(nn.Module):
    """Min-sum message passing layer (message_passing_op_cuda.py:95-178)."""

    @property
    def L1(self):
        return self._L1

    @property
    def L2(self):
        return self._L2

    def setL1L2(self, value_L1, value_L2):
        if value_L1 > 0 and value_L2 > 0 and value_L1 <= value_L2:
            self._L1 = value_L1
            self._L2 = value_L2
        elif value_L1 < 0 or value_L2 < 0:
            raise ValueError("L1 or L2 is < 0!")
        elif value_L1 > value_L2:
            raise ValueError("L1 must be smaller than or equal L2!")

    def __init__(self, device, max_iter, num_labels, delta, mode='min-sum', grad_mode='reference'):
        super(MessagePassing, self).__init__()
        self.device = device
        self.max_iter = max_iter
        if mode != 'min-sum':
            raise ValueError("Unknown message parsing mode " + mode)
        if grad_mode not in ('reference', 'exact'):
            raise ValueError("Unknown gradient mode " + grad_mode)
        self.mode = mode
        self.grad_mode = grad_mode
        self._L1 = nn.Parameter(torch.tensor(0.1, device=device), requires_grad=True)
        self._L2 = nn.Parameter(torch.tensor(2.5, device=device), requires_grad=True)
        self.softmin = TemperatureSoftmin(dim=3, init_temp=1.0, device=device)
        self.delta = delta
        self.rescaleT = None

    def projectL1L2(self):
        self.L2.data = torch.max(self.L1.data, self.L2.data)

    def forward(self, prob_vol, edge_weights, affinities, offset, messages):
        N, H, W, C = prob_vol.shape
        NUM_DIR = 4
        dev = prob_vol.device
        if edge_weights is None:
            edge_weights = torch.ones((N, NUM_DIR, H, W), device=dev)
        if affinities is None:
            if (self._L1 is not None) and (self._L2 is not None):
                affinities = torch.zeros((N, NUM_DIR, 3, H, W), dtype=torch.float, device=dev)
                affinities[:, :, :2, :, :] = self._L1
                affinities[:, :, 2, :, :] = self._L2
        else:
            if (self._L1 is not None and self._L2 is None) or (self._L1 is None and self._L2 is not None):
                raise ValueError("L1 or L2 is None and affinities are not set!")
        if offset is None:
            offset = torch.zeros((N, NUM_DIR, H, W), device=dev)
        if self.mode != 'min-sum':
            raise NotImplementedError("message parsing mode " + self.mode + " is currently not implemented!")
        cost = -prob_vol
        if self.grad_mode == 'exact':
            if self.max_iter > 0:
                messages = LBPMinSumExactChainFunction.apply(cost, edge_weights, messages, self.delta, affinities,
                                                             offset, self.max_iter)
        else:
            for it in range(self.max_iter):
                messages = LBPMinSumFunction.apply(cost, edge_weights, messages, self.delta, affinities, offset)
        beliefs = messages.sum(dim=1) + cost
        beliefs = self.softmin.forward(beliefs)
        return beliefs, messages
