"""Stereo SAD autograd Function (reference ops/sad/stereo_sad.py:5-19)."""
import torch

from ... import pytorch_cuda_stereo_sad_op


class StereoMatchingSadFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, f0, f1, min_disp, max_disp, step=1.0):
        ctx.save_for_backward(f0, f1)
        ctx.min_disp, ctx.max_disp, ctx.step = int(min_disp), int(max_disp), float(step)
        return pytorch_cuda_stereo_sad_op.forward(f0, f1, min_disp, max_disp, step)

    @staticmethod
    def backward(ctx, in_grad):
        f0, f1 = ctx.saved_tensors
        if ctx.step != 1.0:
            raise ValueError("Error: Backward for step != 1 is not implemented!")
        df0, df1 = pytorch_cuda_stereo_sad_op.backward(f0, f1, ctx.min_disp, ctx.max_disp, in_grad.contiguous())
        return df0, df1, None, None, None


class StereoMatchingSadSoftmaxFunction(torch.autograd.Function):
    """prob = softmax(-SAD(f0, f1) / T) over the disparities (step 1), the forward fused into one
    kernel: the composition StereoMatchingSadFunction -> negation -> TemperatureSoftmax of the
    reference's StereoMatchingSad.forward (matching.py:22-25, :45-47; corenet.py:97-104).
    Backward: with z = -SAD / T, dz = p * (g - sum_d g p), dSAD = -dz / T and
    dT = sum(dz * SAD) / T^2 (SAD recomputed), then the SAD backward."""

    @staticmethod
    def forward(ctx, f0, f1, min_disp, max_disp, T):
        t = float(T)
        prob = pytorch_cuda_stereo_sad_op.forward_softmax(f0, f1, min_disp, max_disp, -1.0 / t, 1.0)
        ctx.save_for_backward(f0, f1, prob, T)
        ctx.min_disp, ctx.max_disp = int(min_disp), int(max_disp)
        return prob

    @staticmethod
    def backward(ctx, g):
        f0, f1, p, T = ctx.saved_tensors
        t = float(T)
        g = g.contiguous()
        dz = p * (g - (g * p).sum(dim=-1, keepdim=True))
        sad = pytorch_cuda_stereo_sad_op.forward(f0, f1, ctx.min_disp, ctx.max_disp)
        dT = (dz * sad).sum().reshape(T.shape) / (t * t)
        dsad = (dz * (-1.0 / t)).contiguous()
        df0, df1 = pytorch_cuda_stereo_sad_op.backward(f0, f1, ctx.min_disp, ctx.max_disp, dsad)
        return df0, df1, None, None, dT
