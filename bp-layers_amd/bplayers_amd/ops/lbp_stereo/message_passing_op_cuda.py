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
from .._pairwise import pack_stereo_jump, truncated_linear, unpack_stereo_jump_grad


def construct_pw_tensor(L1, L2, K):
    """[1,1,K,K] P1/P2 energy matrix (0 on the diagonal, L1 one label apart, L2 further;
    message_passing_op_cuda.py:9-20)."""
    return torch.tensor(truncated_linear(K, float(L1), float(L2)), dtype=torch.float)[None, None]


def pack_affinities(affinities, offset):
    """[N,4,K_in,H,W] affinities (K_in = 3: (P1-, P1+, P2); 5: (L2-, L2+, L1-, L1+, L3)) plus
    the [N,4,H,W] window offset -> [N,4,K_in+3,H,W] jump vectors (message_passing_op_cuda.py:35-56;
    slot layout in ops._pairwise.stereo_jump_layout)."""
    return pack_stereo_jump(affinities, offset)


def unpack_affinity_grad(gJ):
    """Adjoint of pack_affinities (message_passing_op_cuda.py:75-91): (grad_affinities,
    grad_offset); the gradient of the constant centre slot is dropped."""
    return unpack_stereo_jump_grad(gJ)


class LBPMinSumFunction(torch.autograd.Function):
    """One LBP iteration with the reference's semantics (in-place messages)."""

    @staticmethod
    def forward(ctx, cost, edge, messages, delta, affinities, offset, msg_zero=False):
        jump = pack_affinities(affinities, offset.to(affinities.device))
        messages, messages_argmin, message_scale = lbp.forward_minsum(cost, jump, edge, messages, delta,
                                                                      msg_in_zero=bool(msg_zero))
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
        return grad_cost, grad_edge, grad_message, None, grad_aff, grad_offset, None


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
    def forward(ctx, cost, edge, messages, delta, affinities, offset, max_iter, msg_zero=False):
        jump = pack_affinities(affinities, offset.to(affinities.device))
        out, recs = _lbp_op.exact_chain_forward(nat.LBP_STEREO, cost, jump, edge, messages.contiguous(), delta,
                                                int(max_iter), msg_in_zero=bool(msg_zero))
        ctx.save_for_backward(cost, jump, edge, *[x for r in recs for x in r])
        ctx.delta = int(delta)
        ctx.iters = int(max_iter)
        return out

    @staticmethod
    def backward(ctx, in_grad):
        cost, jump, edge, *flat = ctx.saved_tensors
        recs = [tuple(flat[3 * i:3 * i + 3]) for i in range(ctx.iters)]
        grad_cost, grad_jump, grad_edge, grad_message = _lbp_op.exact_chain_backward(
            nat.LBP_STEREO, cost, jump, edge, in_grad, recs, delta=ctx.delta,
            want_grad_messages=ctx.needs_input_grad[2])
        grad_aff, grad_offset = unpack_affinity_grad(grad_jump)
        return grad_cost, grad_edge, grad_message, None, grad_aff, grad_offset, None, None


def check_modes(mode, grad_mode):
    """The layers implement min-sum message passing only, with the reference's gradient or the
    exact adjoint."""
    if mode != 'min-sum':
        raise ValueError(f"message passing mode {mode!r} is not supported (only 'min-sum')")
    if grad_mode not in ('reference', 'exact'):
        raise ValueError(f"gradient mode {grad_mode!r} is not supported ('reference' or 'exact')")


class MessagePassing(nn.Module):
    """Min-sum message passing layer (message_passing_op_cuda.py:95-178): ``max_iter`` LBP
    iterations on cost = -prob_vol, then beliefs = softmin(cost + sum of the four messages).
    Without per-pixel affinities the global parameters ``L1`` / ``L2`` (P1 for both
    neighbours, P2) are used for every pixel and direction."""

    def __init__(self, device, max_iter, num_labels, delta, mode='min-sum', grad_mode='reference'):
        super(MessagePassing, self).__init__()
        check_modes(mode, grad_mode)
        self.device = device
        self.max_iter = max_iter
        self.mode = mode
        self.grad_mode = grad_mode
        self.delta = delta
        self.rescaleT = None
        self._L1 = nn.Parameter(torch.tensor(0.1, device=device), requires_grad=True)
        self._L2 = nn.Parameter(torch.tensor(2.5, device=device), requires_grad=True)
        self.softmin = TemperatureSoftmin(dim=3, init_temp=1.0, device=device)

    @property
    def L1(self):
        return self._L1

    @property
    def L2(self):
        return self._L2

    def setL1L2(self, value_L1, value_L2):
        """Replace both parameters when 0 < L1 <= L2; negative values or L1 > L2 raise, and a
        zero leaves the parameters unchanged (the reference's behaviour)."""
        if min(value_L1, value_L2) < 0:
            raise ValueError(f"L1 and L2 must be non-negative (got {value_L1}, {value_L2})")
        if value_L1 > value_L2:
            raise ValueError(f"L1 must not exceed L2 (got {value_L1} > {value_L2})")
        if value_L1 > 0 and value_L2 > 0:
            self._L1, self._L2 = value_L1, value_L2

    def projectL1L2(self):
        self.L2.data = torch.max(self.L1.data, self.L2.data)

    def global_affinities(self, N, H, W, device):
        """(P1, P1, P2) = (L1, L1, L2) at every pixel and direction: [N,4,3,H,W]."""
        aff = torch.zeros((N, 4, 3, H, W), dtype=torch.float, device=device)
        aff[:, :, :2] = self._L1
        aff[:, :, 2] = self._L2
        return aff

    def forward(self, prob_vol, edge_weights, affinities, offset, messages, messages_zero=False):
        """``messages_zero`` (keyword, not in the reference's signature): ``messages`` are all
        zeros (as BP creates them), which the first iteration then need not read."""
        N, H, W, C = prob_vol.shape
        dev = prob_vol.device
        have = (self._L1 is not None, self._L2 is not None)
        if affinities is None and all(have):
            affinities = self.global_affinities(N, H, W, dev)
        elif affinities is not None and any(have) and not all(have):
            raise ValueError("per-pixel affinities were given while only one of L1 / L2 is set")
        edge_weights = torch.ones((N, 4, H, W), device=dev) if edge_weights is None else edge_weights
        offset = torch.zeros((N, 4, H, W), device=dev) if offset is None else offset
        cost = -prob_vol
        if self.grad_mode == 'exact':
            if self.max_iter > 0:
                messages = LBPMinSumExactChainFunction.apply(cost, edge_weights, messages, self.delta, affinities,
                                                             offset, self.max_iter, messages_zero)
        else:
            for it in range(self.max_iter):
                messages = LBPMinSumFunction.apply(cost, edge_weights, messages, self.delta, affinities, offset,
                                                   messages_zero and it == 0)
        beliefs = self.softmin.forward(messages.sum(dim=1) + cost)
        return beliefs, messages
