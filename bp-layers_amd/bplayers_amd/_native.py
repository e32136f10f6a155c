"""ctypes binding of libbpl_amd.so (the C ABI declared in include/bpl_amd.h).

torch is imported first so that the library binds to the HIP runtime torch already
loaded (both use the soname libamdhip64.so.7).  There is no CPU fallback: if the
library is missing every op raises.
"""
import ctypes
import os

import torch

_LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib", "libbpl_amd.so")
_lib = None

# kinds / modes (include/bpl_amd.h)
LBP_STEREO = 0
LBP_SEMANTIC_GLOBAL = 1
LBP_SEMANTIC_PIXEL = 2
GRAD_REFERENCE = 0
GRAD_EXACT = 1
# chained exact iterations: in_grad's / grad_msg's LEFT and RIGHT slots at the row ends only
GRAD_IN_H_ROW_ENDS = 0x10
GRAD_OUT_H_ROW_ENDS = 0x20
GRAD_ACCUMULATE = 0x40
GRAD_OUT_UD_SHARED = 0x80
GRAD_IN_UD_SHARED = 0x100
FWD_MSG_IN_ZERO = 0x1  # BPL_FWD_MSG_IN_ZERO
LBP_BWD_NOALLOC_MAX_BYTES = 128 << 20  # BPL_LBP_BWD_NOALLOC_MAX_BYTES

_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_I = ctypes.c_int
_F = ctypes.c_float

_SIGS = {
    "bpl_version": ([], _I),
    "bpl_last_error": ([], ctypes.c_char_p),
    "bpl_sad_stereo_fwd": ([_P, _P, _P, _I64, _I64, _I64, _I64, _I, _I, _F, _P], _I),
    "bpl_sad_stereo_softmax_fwd": ([_P, _P, _P, _I64, _I64, _I64, _I64, _I, _I, _F, _F, _P], _I),
    "bpl_sad_stereo_bwd": ([_P, _P, _P, _P, _P, _I64, _I64, _I64, _I64, _I, _I, _P], _I),
    "bpl_sad_flow_fwd": ([_P, _P, _P, _P, _P, _P, _I64, _I64, _I64, _I64, _I, _I, _I, _I, _I, _P], _I),
    "bpl_sad_flow_bwd": ([_P, _P, _P, _P, _P, _P, _P, _P, _I64, _I64, _I64, _I64, _I, _P], _I),
    "bpl_lbp_fwd": ([_I, _P, _P, _I64, _I64, _P, _P, _P, _P, _P, _P, _I64, _I64, _I64, _I64, _I, _P], _I),
    "bpl_lbp_fwd_workspace_size": ([_I, _I64, _I64, _I64, _I64], _I64),
    "bpl_lbp_fwd_ws": ([_I, _P, _P, _I64, _I64, _P, _P, _P, _P, _P, _P, _I64, _I64, _I64, _I64, _I, _P, _I64, _I, _P],
                       _I),
    "bpl_lbp_bwd_workspace_size": ([_I, _I64, _I64, _I64, _I64], _I64),
    "bpl_softmax_labels": ([_P, _P, _I64, _I64, _F, _F, _P], _I),
    "bpl_lbp_beliefs": ([_P, _P, _P, _P, _I64, _I64, _I64, _I64, _F, _P], _I),
    "bpl_lbp_bwd": ([_I, _I, _P, _I64, _I64, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I64,
                     _I64, _I64, _I64, _I64, _I, _P], _I),
    "bpl_lbp_bwd_noalloc": ([_I, _I, _P, _I64, _I64, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                             _I64, _I64, _I64, _I64, _I, _P], _I),
    "bpl_timing_enable": ([_I], _I),
    "bpl_timing_reset": ([], _I),
    "bpl_timing_read": ([_I, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(_I64)], _I),
    "bpl_timing_name": ([_I], ctypes.c_char_p),
}
KERNEL_COUNT = 13  # BPL_KERNEL_COUNT


class BplError(RuntimeError):
    """A libbpl_amd entry point returned an error code."""


def library_path():
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(
                f"libbpl_amd.so not found at {_LIB_PATH}; build it with "
                "`python bp-layers_amd/build.py` (hipcc, gfx950)")
        L = ctypes.CDLL(_LIB_PATH)
        for name, (args, res) in _SIGS.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue  # reported by call(); tests/test_capi.py requires every symbol
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def call(name, *args):
    fn = getattr(lib(), name, None)
    if fn is None:
        raise BplError(f"{name} is not exported by {_LIB_PATH}")
    rc = fn(*args)
    if rc != 0:
        msg = lib().bpl_last_error().decode(errors="replace")
        raise BplError(f"{name} failed ({rc}): {msg}")
    return rc


def ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def stream_of(t):
    return ctypes.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)


def check_input(t, name):
    """The reference's CHECK_INPUT (ops/lbp_stereo/src/lbp.cpp:8-10) plus the fp32 it implies."""
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a tensor")
    if not t.is_cuda:
        raise RuntimeError(f"{name} must be a CUDA tensor")
    if not t.is_contiguous():
        raise RuntimeError(f"{name} must be contiguous")
    if t.dtype != torch.float32:
        raise RuntimeError(f"{name} must be float32 (got {t.dtype})")


# ---------------------------------------------------------------- kernel timing
def timing_enable(on=True):
    """Bracket every library kernel launch with HIP events on its stream (or stop doing so)."""
    call("bpl_timing_enable", 1 if on else 0)


def timing_reset():
    call("bpl_timing_reset")


def timing_read():
    """{kernel class name: (total ms, launches)} for the classes launched since the reset;
    waits for the recorded events."""
    out = {}
    for k in range(KERNEL_COUNT):
        ms, n = ctypes.c_double(0.0), _I64(0)
        call("bpl_timing_read", k, ctypes.byref(ms), ctypes.byref(n))
        if n.value:
            out[lib().bpl_timing_name(k).decode()] = (ms.value, n.value)
    return out
