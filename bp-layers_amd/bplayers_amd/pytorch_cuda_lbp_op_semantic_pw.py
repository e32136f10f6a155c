"""Drop-in for the ops/lbp_semantic_pw build of the reference's ``pytorch_cuda_lbp_op`` extension
(ops/lbp_semantic_pw/src/lbp.cpp:16-46), backed by libbpl_amd.  Semantics and deviations: see bplayers_amd._lbp_op."""
from . import _lbp_op
from ._native import LBP_SEMANTIC_GLOBAL as KIND


def forward_minsum(cost, jump, edge, messages, delta, *, out=None):
    """One LBP iteration (RIGHT, LEFT, DOWN, UP); ``messages`` is updated in place unless
    ``out`` is given.  Returns [messages, messages_argmin, message_scale]."""
    return _lbp_op.forward_minsum(KIND, cost, jump, edge, messages, delta, out=out)


def backward_minsum(cost, jump, edge, in_grad, messages, messages_argmin, message_scale, *, mode="reference",
                    delta=None, in_h_row_ends=False, out_h_row_ends=False, grad_messages_out=None):
    """Gradients [grad_unary, grad_pairwise, grad_edge, grad_messages] of one iteration
    (``delta``: the forward's, when known; ``in_h_row_ends`` / ``out_h_row_ends``: chained exact
    iterations, see _lbp_op)."""
    return _lbp_op.backward_minsum(KIND, cost, jump, edge, in_grad, messages, messages_argmin, message_scale, mode=mode,
                                   delta=delta, in_h_row_ends=in_h_row_ends, out_h_row_ends=out_h_row_ends,
                                   grad_messages_out=grad_messages_out)
