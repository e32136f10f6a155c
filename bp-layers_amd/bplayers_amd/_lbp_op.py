"""The min-sum LBP op behind the three ``pytorch_cuda_lbp_op`` modules.

Reference interface (ops/lbp_{stereo,semantic_pw,semantic_pw_pixel}/src/lbp.cpp):
  forward_minsum(cost, jump, edge, messages, delta) -> [messages, messages_argmin, message_scale]
  backward_minsum(cost, jump, edge, in_grad, messages, messages_argmin, message_scale)
      -> [grad_unary, grad_pairwise, grad_edge, grad_messages]
  (the semantic-pixel build's backward has no ``jump`` argument).

Behaviour kept from the reference: inputs must be CUDA, contiguous (and float32);
``messages`` is updated in place and returned; the backward's grad_messages is a copy of
in_grad (the reference semantics, ``mode="reference"``).

Deliberate differences (all internal to the autograd Functions that carry these tensors):
  * messages_argmin is uint8 [N,4,H,W,C]: the winning sender label, or 255 when the value
    came from the truncation term (whose label, the global argmin, is kept per receiver);
  * message_scale is float32 [N,4,H,W] = sum_c exp(msg) as in the reference; its storage
    additionally carries the per-receiver global argmins (uint8) after the floats, which is
    how backward_minsum receives them through the reference's argument list;
  * every batch element is processed; batch-shared pairwise gradients are summed;
  * ``mode="exact"`` (keyword) returns the exact adjoint w.r.t. the input messages;
  * ``out=`` (keyword) writes the messages out of place, leaving ``messages`` untouched;
  * ``msg_in_zero=True`` (keyword) declares ``messages`` all zeros (the first iteration from
    zero messages), which the stereo kernels then need not read;
  * ``in_h_row_ends`` / ``out_h_row_ends`` (keywords, exact mode) chain the backward calls of
    consecutive iterations: in the exact adjoint the LEFT/RIGHT slots of the input-message
    gradient are zero away from the row ends, and its DOWN slot equals its UP slot away from
    rows 0 and H-1, so a call may skip reading those entries in in_grad (``in_h_row_ends``) or
    writing them in its grad_messages (``out_h_row_ends``; they are then left uninitialised and
    the tensor is only fit to be the in_grad of a call with ``in_h_row_ends``);
  * ``grad_messages_out=`` (keyword) receives the input-message gradient; in the reference mode
    it may be ``in_grad`` itself, which then passes through uncopied (for autograd callers).
"""
import torch

from . import _native as nat


def _alloc_scale_and_gargmin(N, H, W, device):
    n = N * 4 * H * W
    buf = torch.empty(n * 5, dtype=torch.uint8, device=device)
    scale = buf[: n * 4].view(torch.float32).view(N, 4, H, W)
    garg = buf[n * 4:].view(N, 4, H, W)
    return scale, garg


def _gargmin_of(message_scale):
    n = message_scale.numel()
    st = message_scale.untyped_storage()
    if message_scale.dtype != torch.float32 or message_scale.storage_offset() != 0 or st.nbytes() < 5 * n:
        raise RuntimeError("message_scale was not produced by this library's forward_minsum")
    g = torch.empty(0, dtype=torch.uint8, device=message_scale.device)
    g.set_(st, 4 * n, (n,))
    return g.view(message_scale.shape)


def _check_common(kind, cost, jump, edge, messages):
    nat.check_input(cost, "cost")
    if jump is not None:
        nat.check_input(jump, "jump")
    nat.check_input(edge, "edge")
    nat.check_input(messages, "messages")
    if cost.dim() != 4:
        raise RuntimeError(f"cost must be [N,H,W,C], got {tuple(cost.shape)}")
    N, H, W, C = cost.shape
    if tuple(messages.shape) != (N, 4, H, W, C):
        raise RuntimeError(f"messages must be {(N, 4, H, W, C)}, got {tuple(messages.shape)}")
    if tuple(edge.shape) != (N, 4, H, W):
        raise RuntimeError(f"edge must be {(N, 4, H, W)}, got {tuple(edge.shape)}")
    for t, nm in ((jump, "jump"), (edge, "edge"), (messages, "messages")):
        if t is not None and t.device != cost.device:
            raise RuntimeError(f"{nm} is on {t.device}, cost on {cost.device}")
    K, Nj = 0, 1
    if kind == nat.LBP_STEREO:
        if jump.dim() != 5 or jump.shape[1] != 4 or tuple(jump.shape[3:]) != (H, W) or jump.shape[0] not in (1, N):
            raise RuntimeError(f"stereo jump must be [N|1,4,K,H,W], got {tuple(jump.shape)}")
        K, Nj = int(jump.shape[2]), int(jump.shape[0])
    elif kind == nat.LBP_SEMANTIC_GLOBAL:
        if tuple(jump.shape[1:]) != (4, C, C) or jump.shape[0] < 1:
            raise RuntimeError(f"semantic jump must be [1,4,C,C], got {tuple(jump.shape)}")
    elif jump is not None:
        if tuple(jump.shape) != (4, H, W, C, C):
            raise RuntimeError(f"pixel jump must be [4,H,W,C,C], got {tuple(jump.shape)}")
    return N, H, W, C, K, Nj


def forward_minsum(kind, cost, jump, edge, messages, delta, out=None, msg_in_zero=False):
    N, H, W, C, K, Nj = _check_common(kind, cost, jump, edge, messages)
    msg_out = messages
    if out is not None:
        nat.check_input(out, "out")
        if out.shape != messages.shape:
            raise RuntimeError("out must have the shape of messages")
        msg_out = out
    argmin = torch.empty((N, 4, H, W, C), dtype=torch.uint8, device=cost.device)
    scale, garg = _alloc_scale_and_gargmin(N, H, W, cost.device)
    delta = max(0, min(int(delta), 65535))
    # stereo iterations of the fast step's shape run with a workspace plane (bpl_lbp_fwd_ws:
    # the vertical phase reads (cost + M_R) + M_L as one plane; the same outputs bit for bit)
    ws = None
    if kind == nat.LBP_STEREO and K == 6 and delta == 3 and C > 16:
        ws = torch.empty(int(nat.lib().bpl_lbp_fwd_workspace_size(kind, N, H, W, C)), dtype=torch.uint8,
                         device=cost.device)
    nat.call("bpl_lbp_fwd_ws", kind, nat.ptr(cost), nat.ptr(jump), K, Nj, nat.ptr(edge), nat.ptr(messages),
             nat.ptr(msg_out), nat.ptr(argmin), nat.ptr(garg), nat.ptr(scale), N, H, W, C, delta,
             nat.ptr(ws), 0 if ws is None else ws.numel(), nat.FWD_MSG_IN_ZERO if msg_in_zero else 0,
             nat.stream_of(cost))
    return [msg_out, argmin, scale]


def backward_minsum(kind, cost, jump, edge, in_grad, messages, messages_argmin, message_scale,
                    mode="reference", delta=None, in_h_row_ends=False, out_h_row_ends=False,
                    grad_messages_out=None, accumulate_into=None, want_grad_messages=True):
    """Backward of one LBP iteration (bpl_lbp_bwd).  ``accumulate_into``: the (grad_cost,
    grad_pairwise, grad_edge) of an earlier call with these sizes, to which this call's
    gradients are added in place (BPL_GRAD_ACCUMULATE) and which are returned.
    ``want_grad_messages=False``: the input-message gradient is not computed (returned as None)."""
    N, H, W, C, K, Nj = _check_common(kind, cost, jump, edge, messages)
    nat.check_input(in_grad, "in_grad")
    if in_grad.shape != messages.shape:
        raise RuntimeError("in_grad must have the shape of messages")
    if messages_argmin.dtype != torch.uint8 or messages_argmin.shape != messages.shape:
        raise RuntimeError("messages_argmin was not produced by this library's forward_minsum")
    garg = _gargmin_of(message_scale)
    m = {"reference": nat.GRAD_REFERENCE, "exact": nat.GRAD_EXACT}[mode]
    if (in_h_row_ends or out_h_row_ends) and mode != "exact":
        raise RuntimeError("in_h_row_ends / out_h_row_ends apply to mode='exact' only")
    # the U/D slots travel with the row-end protocol: equal away from rows 0 and H-1, so a chained
    # gradient carries the DOWN slot at those rows only
    m |= (nat.GRAD_IN_H_ROW_ENDS | nat.GRAD_IN_UD_SHARED if in_h_row_ends else 0) | \
        (nat.GRAD_OUT_H_ROW_ENDS | nat.GRAD_OUT_UD_SHARED if out_h_row_ends else 0)
    dev = cost.device
    grad_cost = torch.empty_like(cost)
    if kind == nat.LBP_STEREO:
        grad_jump = torch.empty_like(jump)
        gj_ret = grad_jump
    elif kind == nat.LBP_SEMANTIC_GLOBAL:
        # reference shape [N,4,C,C] with every batch element accumulated into index 0
        gj_ret = torch.zeros((N, 4, C, C), dtype=torch.float32, device=dev) if N > 1 else None
        grad_jump = gj_ret[:1] if N > 1 else torch.empty((1, 4, C, C), dtype=torch.float32, device=dev)
        gj_ret = gj_ret if N > 1 else grad_jump
    else:
        if mode == "exact" and jump is None:
            raise RuntimeError("the exact backward of the per-pixel variant needs the pairwise tensor (jump=)")
        grad_jump = torch.empty((4, H, W, C, C), dtype=torch.float32, device=dev)
        gj_ret = grad_jump
    grad_edge = torch.empty((N, 4, H, W), dtype=torch.float32, device=dev)
    if accumulate_into is not None:
        if kind == nat.LBP_SEMANTIC_PIXEL:
            raise RuntimeError("accumulate_into applies to the stereo and semantic-global kinds")
        acc = tuple(accumulate_into)
        for t_, ref, name in zip(acc, (grad_cost, gj_ret, grad_edge), ("grad_cost", "grad_pairwise", "grad_edge")):
            nat.check_input(t_, name)
            if t_.shape != ref.shape:
                raise RuntimeError(f"accumulate_into: {name} has shape {tuple(t_.shape)}, expected {tuple(ref.shape)}")
        grad_cost, gj_ret, grad_edge = acc
        grad_jump = gj_ret[:1] if kind == nat.LBP_SEMANTIC_GLOBAL else gj_ret
        m |= nat.GRAD_ACCUMULATE
    if not want_grad_messages:
        grad_msg = None
    elif grad_messages_out is not None:
        nat.check_input(grad_messages_out, "grad_messages_out")
        if grad_messages_out.shape != messages.shape:
            raise RuntimeError("grad_messages_out must have the shape of messages")
        g0, i0, nb = grad_messages_out.data_ptr(), in_grad.data_ptr(), in_grad.numel() * 4
        if g0 == i0 and mode != "reference":
            raise RuntimeError("grad_messages_out may be in_grad in the reference mode only")
        if g0 != i0 and not (g0 + nb <= i0 or i0 + nb <= g0):
            raise RuntimeError("grad_messages_out partly overlaps in_grad")
        grad_msg = grad_messages_out
    else:
        grad_msg = torch.empty_like(messages)
    ws_bytes = int(nat.lib().bpl_lbp_bwd_workspace_size(kind, N, H, W, C))
    ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=dev)
    nat.call("bpl_lbp_bwd", kind, m, nat.ptr(jump), K, Nj, nat.ptr(edge), nat.ptr(in_grad), nat.ptr(messages),
             nat.ptr(messages_argmin), nat.ptr(garg), nat.ptr(grad_cost), nat.ptr(grad_jump),
             nat.ptr(grad_edge), nat.ptr(grad_msg), nat.ptr(ws), ws_bytes, N, H, W, C,
             -1 if delta is None else int(delta), nat.stream_of(cost))
    return [grad_cost, gj_ret, grad_edge, grad_msg]


def exact_chain_forward(kind, cost, jump, edge, messages, delta, iters, msg_in_zero=False):
    """``iters`` iterations from ``messages`` (left untouched), each into fresh messages.
    Returns the final messages and the per-iteration records (messages, argmin, scale) that
    exact_chain_backward replays."""
    recs = []
    m = messages
    for t in range(iters):
        m, am, sc = forward_minsum(kind, cost, jump, edge, m, delta, out=torch.empty_like(m),
                                   msg_in_zero=msg_in_zero and t == 0)
        recs.append((m, am, sc))
    return m, recs


def exact_chain_backward(kind, cost, jump, edge, in_grad, recs, delta=None, want_grad_messages=True):
    """Exact adjoint of exact_chain_forward: the backward calls of consecutive iterations pass
    each other message gradients whose LEFT/RIGHT slots are written and read at the row ends
    only (zero elsewhere in the exact adjoint); the top call reads the full ``in_grad`` and the
    bottom call writes the full gradient w.r.t. the initial messages.  Returns
    (grad_cost, grad_pairwise, grad_edge, grad_messages) with the first three summed over the
    iterations (top iteration first; every later call adds to them in the kernels, except for
    the per-pixel variant, whose pairwise gradient is summed here).  ``want_grad_messages=False``:
    the bottom call computes no initial-message gradient (returned as None)."""
    T = len(recs)
    if T == 0:
        raise RuntimeError("exact_chain_backward needs at least one iteration")
    g = in_grad.contiguous()
    tot = None
    in_kernel = kind != nat.LBP_SEMANTIC_PIXEL
    for t in reversed(range(T)):
        m, am, sc = recs[t]
        gc, gj, ge, g = backward_minsum(kind, cost, jump, edge, g, m, am, sc, mode="exact", delta=delta,
                                        in_h_row_ends=t < T - 1, out_h_row_ends=t > 0,
                                        accumulate_into=tot if in_kernel else None,
                                        want_grad_messages=want_grad_messages or t > 0)
        if tot is None or in_kernel:
            tot = [gc, gj, ge]
        else:
            tot[0] += gc
            tot[1] += gj
            tot[2] += ge
    return tot[0], tot[1], tot[2], g
