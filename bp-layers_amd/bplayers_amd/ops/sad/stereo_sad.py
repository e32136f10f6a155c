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
