"""Semantic min-sum LBP with a per-pixel C x C pairwise matrix per direction
(reference ops/lbp_semantic_pw_pixel/message_passing_op_pw_pixel.py:23-109)."""
import torch
import torch.nn as nn

from ... import _lbp_op
from ... import _native as nat
from ... import pytorch_cuda_lbp_op_semantic_pw_pixel as lbp
from ...corenet import TemperatureSoftmin
from ..lbp_stereo.message_passing_op_cuda import check_modes, construct_pw_tensor  # noqa: F401


def shift_pairwise(jump):
    """[2,H,W,C,C] (horizontal, vertical) -> [4,H,W,C,C]: LEFT = jump[0]; RIGHT at x =
    jump[0] at x-1 transposed (0 at x=0); UP = jump[1]; DOWN at y = jump[1] at y-1
    transposed (0 at y=0) (message_passing_op_pw_pixel.py:27-38)."""
    s = torch.zeros((jump.shape[0] + 2,) + tuple(jump.shape[1:]), dtype=jump.dtype, device=jump.device)
    s[0] = jump[0]
    s[1, :, 1:] = jump[0, :, :-1].permute((0, 1, 3, 2))
    s[2] = jump[1]
    s[3, 1:] = jump[1, :-1].permute((0, 1, 3, 2))
    return s.contiguous()


def fold_pairwise_grad(grad_jump):
    """Adjoint of shift_pairwise (message_passing_op_pw_pixel.py:55-58)."""
    g0 = grad_jump[0].clone()
    g0[:, :-1] += grad_jump[1, :, 1:].permute((0, 1, 3, 2))
    g1 = grad_jump[2].clone()
    g1[:-1] += grad_jump[3, 1:].permute((0, 1, 3, 2))
    return torch.stack([g0, g1])


class LBPMinSumFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, cost, L1, L2, edge, messages, delta, jump, grad_mode='reference'):
        j = shift_pairwise(jump)
        out = None if grad_mode == 'reference' else torch.empty_like(messages)
        messages, messages_argmin, message_scale = lbp.forward_minsum(cost, j, edge, messages.contiguous(), delta,
                                                                      out=out)
        ctx.save_for_backward(cost, edge, messages, messages_argmin, message_scale, j)
        ctx.grad_mode = grad_mode
        return messages

    @staticmethod
    def backward(ctx, in_grad):
        cost, edge, messages, messages_argmin, message_scale, j = ctx.saved_tensors
        g = in_grad.contiguous()
        # reference mode: grad_messages = in_grad passes through without a copy
        grad_cost, grad_jump, grad_edge, grad_message = lbp.backward_minsum(
            cost, edge, g, messages, messages_argmin, message_scale, mode=ctx.grad_mode,
            jump=j if ctx.grad_mode == 'exact' else None,
            grad_messages_out=g if ctx.grad_mode == 'reference' else None)
        return grad_cost, None, None, grad_edge, grad_message, None, fold_pairwise_grad(grad_jump), None


class LBPMinSumExactChainFunction(torch.autograd.Function):
    """``max_iter`` (>= 1) iterations with the exact adjoint as one autograd node (the values of
    ``max_iter`` exact-mode LBPMinSumFunction applications; the iterations' backward calls are
    chained through bpl_lbp_bwd's row-end flags)."""

    @staticmethod
    def forward(ctx, cost, edge, messages, delta, jump, max_iter):
        j = shift_pairwise(jump)
        out, recs = _lbp_op.exact_chain_forward(nat.LBP_SEMANTIC_PIXEL, cost, j, edge, messages.contiguous(), delta,
                                                int(max_iter))
        ctx.save_for_backward(cost, j, edge, *[x for r in recs for x in r])
        ctx.iters = int(max_iter)
        return out

    @staticmethod
    def backward(ctx, in_grad):
        cost, j, edge, *flat = ctx.saved_tensors
        recs = [tuple(flat[3 * i:3 * i + 3]) for i in range(ctx.iters)]
        grad_cost, grad_jump, grad_edge, grad_message = _lbp_op.exact_chain_backward(
            nat.LBP_SEMANTIC_PIXEL, cost, j, edge, in_grad, recs)
        return grad_cost, grad_edge, grad_message, None, fold_pairwise_grad(grad_jump), None


class MessagePassing(nn.Module):
    def __init__(self, device, max_iter, num_labels, delta, mode='min-sum', grad_mode='reference'):
        super(MessagePassing, self).__init__()
        check_modes(mode, grad_mode)
        self.device = device
        self.max_iter = max_iter
        self.mode = mode
        self.grad_mode = grad_mode
        self.delta = delta
        self.rescaleT = None
        self.L1 = nn.Parameter(torch.tensor(0.1, device=device), requires_grad=True)
        self.L2 = nn.Parameter(torch.tensor(2.5, device=device), requires_grad=True)
        self.softmin = TemperatureSoftmin(dim=3, init_temp=1.0, device=device)

    def projectL1L2(self):
        self.L2.data = torch.max(self.L1.data, self.L2.data)

    def forward(self, prob_vol, edge_weights, messages, jump):
        N, H, W, C = prob_vol.shape
        edge_weights = torch.ones((N, 4, H, W), device=prob_vol.device) if edge_weights is None else edge_weights
        cost = -prob_vol
        if self.grad_mode == 'exact':
            if self.max_iter > 0:
                messages = LBPMinSumExactChainFunction.apply(cost, edge_weights, messages, self.delta, jump,
                                                             self.max_iter)
        else:
            for _ in range(self.max_iter):
                messages = LBPMinSumFunction.apply(cost, self.L1, self.L2, edge_weights, messages, self.delta, jump,
                                                   self.grad_mode)
        beliefs = messages.sum(dim=1) + cost
        return self.softmin.forward(beliefs)
