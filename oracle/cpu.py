"""ctypes loader of the C++/OpenMP CPU restatement oracle/c/lbp_cpu.cpp (TEST INFRASTRUCTURE).

Builds it with ``make -C oracle`` on first use when the library is missing."""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "c", "libbpl_cpu.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = ctypes.CDLL(LIB)
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def lbp_forward(cost, jump, edge, messages, delta, dtype=np.float32, threads=0):
    """One iteration; returns (messages, argmin labels, from-truncation flags)."""
    cost = np.ascontiguousarray(cost, dtype=dtype)
    jump = np.ascontiguousarray(jump, dtype=dtype)
    edge = np.ascontiguousarray(edge, dtype=dtype)
    min_ = np.ascontiguousarray(messages, dtype=dtype)
    N, H, W, C = cost.shape
    out = np.empty_like(min_)
    am = np.zeros(min_.shape, dtype=np.int32)
    fg = np.zeros(min_.shape, dtype=np.uint8)
    fn = lib().cpu_lbp_fwd_f32 if dtype == np.float32 else lib().cpu_lbp_fwd_f64
    fn(_p(cost), _p(jump), jump.shape[2], jump.shape[0], _p(edge), _p(min_), _p(out), _p(am), _p(fg), N, H, W, C,
       int(delta), int(threads))
    return out, am, fg


def lbp_backward_exact(jump, edge, G, M, argmin, fromg, delta, dtype=np.float32, threads=0):
    jump = np.ascontiguousarray(jump, dtype=dtype)
    edge = np.ascontiguousarray(edge, dtype=dtype)
    G = np.ascontiguousarray(G, dtype=dtype)
    M = np.ascontiguousarray(M, dtype=dtype)
    argmin = np.ascontiguousarray(argmin, dtype=np.int32)
    fromg = np.ascontiguousarray(fromg, dtype=np.uint8)
    N, _, H, W, C = M.shape
    gc = np.empty((N, H, W, C), dtype=dtype)
    gj = np.empty_like(jump)
    ge = np.empty((N, 4, H, W), dtype=dtype)
    gm = np.empty_like(M)
    work = np.empty((4, N, H, W, C), dtype=dtype)
    fn = lib().cpu_lbp_bwd_f32 if dtype == np.float32 else lib().cpu_lbp_bwd_f64
    fn(_p(jump), jump.shape[2], jump.shape[0], _p(edge), _p(G), _p(M), _p(argmin), _p(fromg), _p(gc), _p(gj),
       _p(ge), _p(gm), _p(work), N, H, W, C, int(delta), int(threads))
    return gc, gj, ge, gm


def sad_forward(f0, f1, min_disp, max_disp, threads=0):
    f0 = np.ascontiguousarray(f0, dtype=np.float32)
    f1 = np.ascontiguousarray(f1, dtype=np.float32)
    N, C, H, W = f0.shape
    D = max_disp - min_disp + 1
    out = np.empty((N, H, W, D), dtype=np.float32)
    lib().cpu_sad_fwd_f32(_p(f0), _p(f1), _p(out), N, C, H, W, int(min_disp), D, int(threads))
    return out


def max_threads():
    return int(lib().cpu_max_threads())
