"""Drop-in for the reference's ``pytorch_cuda_stereo_sad_op`` extension
(ops/sad/src/stereo_sad.cpp:10-30), backed by libbpl_amd."""
import torch

from . import _native as nat


def _check(f0, f1):
    nat.check_input(f0, "f0")
    nat.check_input(f1, "f1")
    if f0.dim() != 4 or f0.shape != f1.shape:
        raise RuntimeError(f"f0 and f1 must be [N,C,H,W] of one shape, got {tuple(f0.shape)} and {tuple(f1.shape)}")


def num_disparities(min_disp, max_disp, step=1.0):
    """D = (max_disp - min_disp + 1) / step, truncated (stereo_sad_kernel.cu:122)."""
    return int(float(torch.tensor(float(max_disp - min_disp + 1), dtype=torch.float32) /
                     torch.tensor(float(step), dtype=torch.float32)))


def forward(f0, f1, min_disp, max_disp, step=1.0):
    """cost[n,y,x,i] = sum_c |f0[n,c,y,x] - f1[n,c,y,x-d]|, d = int(i*step)+min_disp; 40 outside."""
    _check(f0, f1)
    N, C, H, W = f0.shape
    D = num_disparities(min_disp, max_disp, step)
    if D < 1:
        raise RuntimeError("empty disparity range")
    cost = torch.empty((N, H, W, D), dtype=torch.float32, device=f0.device)
    nat.call("bpl_sad_stereo_fwd", nat.ptr(f0), nat.ptr(f1), nat.ptr(cost), N, C, H, W, int(min_disp),
             int(max_disp), float(step), nat.stream_of(f0))
    return cost


def forward_softmax(f0, f1, min_disp, max_disp, scale, sign=1.0):
    """sign * softmax(scale * cost) over the disparities, cost = forward(f0, f1, min_disp,
    max_disp) at step 1, in one kernel: the matching epilogue of matching.py:22-25 / :45-47
    with corenet.py:97-104 (scale = -1/T).  An addition to the reference's extension API; the
    result equals forward() followed by the label softmax bit for bit.  D <= 256."""
    _check(f0, f1)
    N, C, H, W = f0.shape
    D = int(max_disp) - int(min_disp) + 1
    if D < 1:
        raise RuntimeError("empty disparity range")
    y = torch.empty((N, H, W, D), dtype=torch.float32, device=f0.device)
    nat.call("bpl_sad_stereo_softmax_fwd", nat.ptr(f0), nat.ptr(f1), nat.ptr(y), N, C, H, W, int(min_disp),
             int(max_disp), float(scale), float(sign), nat.stream_of(f0))
    return y


def backward(f0, f1, min_disp, max_disp, in_grad):
    """[df0, df1] of forward at step 1 (stereo_sad_kernel.cu:64-108)."""
    _check(f0, f1)
    nat.check_input(in_grad, "in_grad")
    N, C, H, W = f0.shape
    D = int(max_disp) - int(min_disp) + 1
    if tuple(in_grad.shape) != (N, H, W, D):
        raise RuntimeError(f"in_grad must be {(N, H, W, D)}, got {tuple(in_grad.shape)}")
    df0 = torch.empty_like(f0)
    df1 = torch.empty_like(f1)
    nat.call("bpl_sad_stereo_bwd", nat.ptr(f0), nat.ptr(f1), nat.ptr(in_grad), nat.ptr(df0), nat.ptr(df1),
             N, C, H, W, int(min_disp), int(max_disp), nat.stream_of(f0))
    return [df0, df1]
