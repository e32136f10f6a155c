"""Flow SAD + min-projection autograd Function (reference ops/flow_mp_sad/flow_mp_sad.py:5-119).

Search windows up to sws=108 are one window, as in the reference.  sws=216 is computed
as one 217x217 window here; the reference tiles it into four 109x109 windows and merges
them with a strict '<' (first tile wins), which gives the same minima and first-index
argmins (checked against the oracle in tests/test_oracle_sad.py).  Other sizes raise, as
in the reference (its sws=432 branch is unfinished)."""
import torch

from ... import pytorch_cuda_flow_mp_sad_op


class FlowMpSadFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, f0, f1, sws):
        if not (sws <= 108 or sws == 2 * 108):
            raise ValueError(f"search window sws={sws} is not supported (sws <= 108 or sws = 216)")
        cv_u, cv_v, u_star, v_star = pytorch_cuda_flow_mp_sad_op.forward(f0, f1, sws, 0, 0, 0, 0)
        ctx.save_for_backward(f0, f1, u_star, v_star)
        ctx.sws = int(sws)
        ctx.mark_non_differentiable(u_star, v_star)
        return cv_u, cv_v, u_star, v_star

    @staticmethod
    def backward(ctx, in_grad_u, in_grad_v, u_star_grad, v_star_grad):
        f0, f1, u_star, v_star = ctx.saved_tensors
        if in_grad_u is None:
            in_grad_u = torch.zeros_like(u_star)
        if in_grad_v is None:
            in_grad_v = torch.zeros_like(v_star)
        df0, df1 = pytorch_cuda_flow_mp_sad_op.backward(f0, f1, ctx.sws, in_grad_u.contiguous(),
                                                        in_grad_v.contiguous(), u_star, v_star)
        return df0, df1, None
