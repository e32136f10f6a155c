"""File formats around the ops: optical-flow ``.flo`` files (reference data.py:47-59), PFM
disparity maps (the output of main_stereo.py:66-67) and the reference's ``.cpt`` parameter
checkpoints (networks.py:41-54 with the per-module hooks, e.g. bp_op_cuda.py:221-237).

Checkpoints are read with ``torch.load(..., weights_only=True)``: a plain state_dict of
tensors loads, anything that would need code from the file is refused."""
import struct

import numpy as np
import torch

FLO_TAG = b"PIEH"  # 202021.25 as float32


def read_flo(path):
    """[H, W, 2] float32 (u, v) from a Middlebury ``.flo`` file: 'PIEH', int32 width, int32
    height, then H*W*2 little-endian float32 (data.py:47-59)."""
    with open(path, "rb") as f:
        if f.read(4) != FLO_TAG:
            raise ValueError("Flow file header does not contain PIEH")
        w, h = np.frombuffer(f.read(8), dtype="<i4")
        data = np.frombuffer(f.read(int(w) * int(h) * 2 * 4), dtype="<f4")
    if data.size != int(w) * int(h) * 2:
        raise ValueError(f"truncated flow file: expected {int(w) * int(h) * 2} values, got {data.size}")
    return data.reshape(int(h), int(w), 2).astype(np.float32)


def write_flo(path, flow):
    """Inverse of read_flo; ``flow`` is [H, W, 2]."""
    flow = np.asarray(flow, dtype="<f4")
    if flow.ndim != 3 or flow.shape[2] != 2:
        raise ValueError(f"flow must be [H, W, 2], got {flow.shape}")
    h, w = flow.shape[:2]
    with open(path, "wb") as f:
        f.write(FLO_TAG)
        f.write(np.array([w, h], dtype="<i4").tobytes())
        f.write(np.ascontiguousarray(flow).tobytes())


def write_pfm(path, image):
    """Single-channel PFM of an [H, W] map with row 0 at the top of the image: header
    'Pf\\n<W> <H>\\n-1.0\\n' (negative scale = little-endian), then the rows bottom to top as the
    format stores them -- the file main_stereo.py writes for a disparity map."""
    img = np.asarray(image, dtype="<f4")
    if img.ndim != 2:
        raise ValueError(f"image must be [H, W], got {img.shape}")
    h, w = img.shape
    with open(path, "wb") as f:
        f.write(b"Pf\n%d %d\n-1.0\n" % (w, h))
        f.write(np.ascontiguousarray(img[::-1]).tobytes())


def read_pfm(path):
    """[H, W] (``Pf``) or [H, W, 3] (``PF``) float32 with row 0 at the top."""
    with open(path, "rb") as f:
        kind = f.readline().strip()
        if kind not in (b"Pf", b"PF"):
            raise ValueError(f"not a PFM file: {kind!r}")
        w, h = (int(v) for v in f.readline().split())
        scale = float(f.readline().strip())
        ch = 3 if kind == b"PF" else 1
        data = np.frombuffer(f.read(w * h * ch * 4), dtype="<f4" if scale < 0 else ">f4")
    if data.size != w * h * ch:
        raise ValueError("truncated PFM file")
    img = data.reshape((h, w, ch) if ch == 3 else (h, w))[::-1]
    return np.ascontiguousarray(img, dtype=np.float32)


def load_checkpoint(module, path, map_location=None):
    """Load a reference ``.cpt`` state_dict into ``module``: weights-only read, then the
    module's ``hook_adjust_checkpoint`` if it has one (networks.py:41-54)."""
    ckpt = torch.load(path, map_location=map_location, weights_only=True)
    if not isinstance(ckpt, dict):
        raise ValueError(f"{path}: not a state_dict")
    hook = getattr(module, "hook_adjust_checkpoint", None)
    if hook is not None:
        ckpt = hook(ckpt)
    module.load_state_dict(ckpt)
    return module
