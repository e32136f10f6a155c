"""Semantic min-sum LBP with a global C x C pairwise matrix per direction
(reference ops/lbp_semantic_pw/message_passing_op_cuda.py:23-120)."""
import torch
import torch.nn as nn

from ... import _lbp_op
from ... import _native as nat
from ... import pytorch_cuda_lbp_op_semantic_pw as lbp
from ...corenet import TemperatureSoftmin
from ..lbp_stereo.message_passing_op_cuda import check_modes, construct_pw_tensor  # noqa: F401


def pack_pairwise(jump, single_pw):
    """pw_mat [1,2,C,C] -> jump_in [1,4,C,C]: LEFT=M_h, RIGHT=M_h^T, UP=M_v, DOWN=M_v^T; with
    single_pw all four are jump[0,0] (message_passing_op_cuda.py:27-40)."""
    if not bool(single_pw):
        j = torch.stack([jump[0, 0], jump[0, 0].t(), jump[0, 1], jump[0, 1].t()])[None]
    else:
        j = torch.stack([jump[0, 0]] * 4)[None]
    return j.contiguous()


def unpack_pairwise_grad(grad_jump, single_pw):
    """Adjoint of pack_pairwise (message_passing_op_cuda.py:57-70)."""
    g = grad_jump[0]
    if not bool(single_pw):
        return torch.stack([g[0] + g[1].t(), g[2] + g[3].t()])[None]
    return (g[0] + g[1] + g[2] + g[3])[None, None]


class LBPMinSumFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, cost, L1, L2, edge, messages, delta, jump, single_pw, grad_mode='reference'):
        jump_in = pack_pairwise(jump, single_pw)
        out = None if grad_mode == 'reference' else torch.empty_like(messages)
        messages, messages_argmin, message_scale = lbp.forward_minsum(cost, jump_in, edge, messages.contiguous(),
                                                                      delta, out=out)
        ctx.save_for_backward(cost, jump_in, edge, messages, messages_argmin, message_scale)
        ctx.single_pw = bool(single_pw)
        ctx.grad_mode = grad_mode
        ctx.delta = int(delta)
        return messages

    @staticmethod
    def backward(ctx, in_grad):
        cost, jump, edge, messages, messages_argmin, message_scale = ctx.saved_tensors
        g = in_grad.contiguous()
        # reference mode: grad_messages = in_grad passes through without a copy
        grad_cost, grad_jump, grad_edge, grad_message = lbp.backward_minsum(
            cost, jump.contiguous(), edge, g, messages, messages_argmin, message_scale,
            mode=ctx.grad_mode, delta=ctx.delta, grad_messages_out=g if ctx.grad_mode == 'reference' else None)
        return (grad_cost, None, None, grad_edge, grad_message, None, unpack_pairwise_grad(grad_jump, ctx.single_pw),
                None, None)


class LBPMinSumExactChainFunction(torch.autograd.Function):
    """``max_iter`` (>= 1) iterations with the exact adjoint as one autograd node (the values of
    ``max_iter`` exact-mode LBPMinSumFunction applications; the iterations' backward calls are
    chained through bpl_lbp_bwd's row-end flags)."""

    @staticmethod
    def forward(ctx, cost, edge, messages, delta, jump, single_pw, max_iter):
        jump_in = pack_pairwise(jump, single_pw)
        out, recs = _lbp_op.exact_chain_forward(nat.LBP_SEMANTIC_GLOBAL, cost, jump_in, edge, messages.contiguous(),
                                                delta, int(max_iter))
        ctx.save_for_backward(cost, jump_in, edge, *[x for r in recs for x in r])
        ctx.single_pw = bool(single_pw)
        ctx.iters = int(max_iter)
        ctx.delta = int(delta)
        return out

    @staticmethod
    def backward(ctx, in_grad):
        cost, jump_in, edge, *flat = ctx.saved_tensors
        recs = [tuple(flat[3 * i:3 * i + 3]) for i in range(ctx.iters)]
        grad_cost, grad_jump, grad_edge, grad_message = _lbp_op.exact_chain_backward(
            nat.LBP_SEMANTIC_GLOBAL, cost, jump_in, edge, in_grad, recs, delta=ctx.delta)
        return grad_cost, grad_edge, grad_message, None, unpack_pairwise_grad(grad_jump, ctx.single_pw), None, None


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

    def forward(self, prob_vol, edge_weights, messages, jump, single_pw=False):
        N, H, W, C = prob_vol.shape
        edge_weights = torch.ones((N, 4, H, W), device=prob_vol.device) if edge_weights is None else edge_weights
        cost = -prob_vol
        if self.grad_mode == 'exact':
            if self.max_iter > 0:
                messages = LBPMinSumExactChainFunction.apply(cost, edge_weights, messages, self.delta, jump,
                                                             bool(single_pw), self.max_iter)
        else:
            for _ in range(self.max_iter):
                messages = LBPMinSumFunction.apply(cost, self.L1, self.L2, edge_weights, messages, self.delta, jump,
                                                   torch.tensor(single_pw), self.grad_mode)
        beliefs = messages.sum(dim=1) + cost
        return self.softmin.forward(beliefs)
