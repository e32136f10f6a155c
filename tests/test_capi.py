"""CPU checks of the C ABI: the library loads, exports every function include/bpl_amd.h
declares, and rejects bad arguments with error codes before touching a GPU."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "bpl_amd.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(bpl_\w+)\s*\(", text, flags=re.M)))


@pytest.fixture(scope="module")
def lib():
    from bplayers_amd import _native
    if not os.path.exists(_native.library_path()):
        import importlib.util
        spec = importlib.util.spec_from_file_location("bpl_build", os.path.join(ROOT, "bp-layers_amd", "build.py"))
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        mod.build()
    return _native.lib()


def test_header_declares_the_expected_surface():
    names = declared_functions()
    for n in ("bpl_version", "bpl_last_error", "bpl_sad_stereo_fwd", "bpl_sad_stereo_bwd", "bpl_sad_flow_fwd",
              "bpl_sad_flow_bwd", "bpl_lbp_fwd", "bpl_lbp_bwd", "bpl_lbp_bwd_workspace_size"):
        assert n in names


def test_library_exports_every_declared_symbol(lib):
    missing = [n for n in declared_functions() if getattr(lib, n, None) is None]
    assert not missing, f"not exported: {missing}"


def test_version_and_error_string(lib):
    assert lib.bpl_version() == 1
    assert isinstance(lib.bpl_last_error(), bytes)


def test_argument_validation_returns_codes(lib):
    from bplayers_amd import _native as nat
    P = ctypes.c_void_p(1)
    rc = lib.bpl_lbp_fwd(7, P, P, 6, 1, P, P, P, P, P, P, 1, 4, 4, 8, 3, None)
    assert rc == -1 and b"unknown kind" in lib.bpl_last_error()
    rc = lib.bpl_lbp_fwd(nat.LBP_STEREO, P, P, 6, 1, P, P, P, P, P, P, 1, 4, 4, 300, 3, None)
    assert rc == -2 and b"exceeds" in lib.bpl_last_error()
    rc = lib.bpl_lbp_fwd(nat.LBP_STEREO, P, P, 20, 1, P, P, P, P, P, P, 1, 4, 4, 8, 3, None)
    assert rc == -2
    rc = lib.bpl_lbp_fwd(nat.LBP_STEREO, P, P, 6, 1, P, P, P, P, None, P, 1, 4, 4, 8, 3, None)
    assert rc == -1 and b"gargmin" in lib.bpl_last_error()
    rc = lib.bpl_lbp_fwd(nat.LBP_STEREO, P, None, 6, 1, P, P, P, P, P, P, 1, 4, 4, 8, 3, None)
    assert rc == -1 and b"null" in lib.bpl_last_error()
    rc = lib.bpl_lbp_bwd(nat.LBP_STEREO, 5, P, 6, 1, P, P, P, P, P, P, P, P, P, P, 1 << 40, 1, 4, 4, 8, 3, None)
    assert rc == -1 and b"mode" in lib.bpl_last_error()
    # the chained-iteration row-end flags belong to the exact mode only
    for fl in (nat.GRAD_IN_H_ROW_ENDS, nat.GRAD_OUT_H_ROW_ENDS, nat.GRAD_IN_UD_SHARED, nat.GRAD_OUT_UD_SHARED):
        rc = lib.bpl_lbp_bwd(nat.LBP_STEREO, nat.GRAD_REFERENCE | fl, P, 6, 1, P, P, P, P, P, P, P, P, P, P,
                             1 << 40, 1, 4, 4, 8, 3, None)
        assert rc == -1 and b"exact mode only" in lib.bpl_last_error()
    rc = lib.bpl_lbp_bwd(nat.LBP_STEREO, nat.GRAD_EXACT | 0x200, P, 6, 1, P, P, P, P, P, P, P, P, P, P,
                         1 << 40, 1, 4, 4, 8, 3, None)
    assert rc == -1 and b"unknown mode" in lib.bpl_last_error()
    # accumulation across iterations: stereo and semantic-global kinds only
    rc = lib.bpl_lbp_bwd(nat.LBP_SEMANTIC_PIXEL, nat.GRAD_EXACT | nat.GRAD_ACCUMULATE, P, 6, 1, P, P, P, P, P,
                         P, P, P, P, P, 1 << 40, 1, 4, 4, 8, 3, None)
    assert rc == -1 and b"ACCUMULATE" in lib.bpl_last_error()
    # grad_msg may alias in_grad only in the reference mode (arguments: ..., in_grad, ..., grad_msg)
    G = ctypes.c_void_p(4096)
    rc = lib.bpl_lbp_bwd(nat.LBP_STEREO, nat.GRAD_EXACT, P, 6, 1, P, G, P, P, P, P, P, P, G, P,
                         1 << 40, 1, 4, 4, 8, 3, None)
    assert rc == -1 and b"alias" in lib.bpl_last_error()
    need = lib.bpl_lbp_bwd_workspace_size(nat.LBP_STEREO, 1, 4, 4, 8)
    # four h-gradient planes, then the accumulation copies of grad_jump (up to 15 entries per
    # direction) and grad_edge
    assert need == (4 * 1 * 4 * 4 * 8 + 1 * 4 * 15 * 4 * 4 + 1 * 4 * 4 * 4) * 4
    rc = lib.bpl_lbp_bwd(nat.LBP_STEREO, 1, P, 6, 1, P, P, P, P, P, P, P, P, G, P, need - 1, 1, 4, 4, 8, 3, None)
    assert rc == -1 and b"workspace" in lib.bpl_last_error()
    rc = lib.bpl_sad_stereo_fwd(P, P, P, 1, 3, 4, 4, 5, 2, 1.0, None)
    assert rc == -1 and b"empty" in lib.bpl_last_error()
    rc = lib.bpl_sad_flow_fwd(P, P, P, P, P, P, 1, 3, 4, 4, -2, 0, 0, 0, 0, None)
    assert rc == -1
    f = ctypes.c_float
    rc = lib.bpl_sad_stereo_softmax_fwd(P, P, P, 1, 3, 4, 4, 0, 256, f(-1.0), f(1.0), None)
    assert rc == -2 and b"exceeds 256" in lib.bpl_last_error()
    rc = lib.bpl_sad_stereo_softmax_fwd(P, None, P, 1, 3, 4, 4, 0, 63, f(-1.0), f(1.0), None)
    assert rc == -1 and b"null" in lib.bpl_last_error()
    rc = lib.bpl_sad_stereo_softmax_fwd(P, P, P, 1, 3, 4, 4, 5, 2, f(-1.0), f(1.0), None)
    assert rc == -1 and b"empty" in lib.bpl_last_error()


def test_python_ops_require_cuda_tensors(lib):
    import torch
    from bplayers_amd import pytorch_cuda_lbp_op_stereo as st
    from bplayers_amd import pytorch_cuda_stereo_sad_op as sad
    x = torch.zeros((1, 4, 4, 8))
    with pytest.raises(RuntimeError, match="CUDA"):
        st.forward_minsum(x, torch.zeros((1, 4, 6, 4, 4)), torch.zeros((1, 4, 4, 4)), torch.zeros((1, 4, 4, 4, 8)), 3)
    with pytest.raises(RuntimeError, match="CUDA"):
        sad.forward(torch.zeros((1, 2, 4, 4)), torch.zeros((1, 2, 4, 4)), 0, 3)


def test_reference_aliases():
    import sys
    import bplayers_amd
    bplayers_amd.install_reference_aliases("semantic_pw")
    import pytorch_cuda_lbp_op
    assert pytorch_cuda_lbp_op.__name__.endswith("semantic_pw")
    assert hasattr(pytorch_cuda_lbp_op, "forward_minsum") and hasattr(pytorch_cuda_lbp_op, "backward_minsum")
    for k in ("pytorch_cuda_lbp_op", "pytorch_cuda_stereo_sad_op", "pytorch_cuda_flow_mp_sad_op"):
        sys.modules.pop(k)


def test_timing_instrumentation_without_launches(lib):
    """Enabling, resetting and reading the timing totals needs no GPU and reports nothing
    before any launch."""
    from bplayers_amd import _native as nat
    nat.timing_enable(True)
    nat.timing_reset()
    assert nat.timing_read() == {}
    nat.timing_enable(False)
    names = [lib.bpl_timing_name(k).decode() for k in range(nat.KERNEL_COUNT)]
    assert len(set(names)) == nat.KERNEL_COUNT and all(names)
    assert lib.bpl_timing_name(nat.KERNEL_COUNT) == b""
    ms = ctypes.c_double(0.0)
    n = ctypes.c_int64(0)
    assert lib.bpl_timing_read(99, ctypes.byref(ms), ctypes.byref(n)) == -1
