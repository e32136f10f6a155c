"""Drop-in for the reference's ``pytorch_cuda_flow_mp_sad_op`` extension
(ops/flow_mp_sad/src/flow_mp_sad.cpp:10-36), backed by libbpl_amd."""
import torch

from . import _native as nat


def _check(f0, f1):
    nat.check_input(f0, "f0")
    nat.check_input(f1, "f1")
    if f0.dim() != 4 or f0.shape != f1.shape:
        raise RuntimeError(f"f0 and f1 must be [N,C,H,W] of one shape, got {tuple(f0.shape)} and {tuple(f1.shape)}")


def forward(f0, f1, sws, offset_u, offset_v, blockIdx_u, blockIdx_v):
    """[cv_u, cv_v, u_star, v_star], each [N,H,W,sws+1]; argmins as float label ids."""
    _check(f0, f1)
    N, C, H, W = f0.shape
    K = int(sws) + 1
    outs = [torch.empty((N, H, W, K), dtype=torch.float32, device=f0.device) for _ in range(4)]
    nat.call("bpl_sad_flow_fwd", nat.ptr(f0), nat.ptr(f1), *[nat.ptr(o) for o in outs], N, C, H, W, int(sws),
             int(offset_u), int(offset_v), int(blockIdx_u), int(blockIdx_v), nat.stream_of(f0))
    return outs


def backward(f0, f1, sws, in_grad_u, in_grad_v, u_star, v_star):
    """[df0, df1] routed through the stored argmins (flow_mp_sad_kernel.cu:126-200)."""
    _check(f0, f1)
    for t, nm in ((in_grad_u, "in_grad_u"), (in_grad_v, "in_grad_v"), (u_star, "u_star"), (v_star, "v_star")):
        nat.check_input(t, nm)
    N, C, H, W = f0.shape
    K = int(sws) + 1
    for t in (in_grad_u, in_grad_v, u_star, v_star):
        if tuple(t.shape) != (N, H, W, K):
            raise RuntimeError(f"flow volumes must be {(N, H, W, K)}, got {tuple(t.shape)}")
    df0 = torch.empty_like(f0)
    df1 = torch.empty_like(f1)
    nat.call("bpl_sad_flow_bwd", nat.ptr(f0), nat.ptr(f1), nat.ptr(in_grad_u), nat.ptr(in_grad_v),
             nat.ptr(u_star), nat.ptr(v_star), nat.ptr(df0), nat.ptr(df1), N, C, H, W, int(sws), nat.stream_of(f0))
    return [df0, df1]
