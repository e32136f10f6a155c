"""GPU test of the kernel-timing instrumentation of libbpl_amd."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def test_timing_books_each_launch_to_its_kernel_class():
    from bplayers_amd import _native as nat
    from bplayers_amd import pytorch_cuda_lbp_op_stereo as st
    from test_gpu_lbp_fwd import make_case
    cost, jump, edge, M0 = make_case("stereo", 1, 6, 9, 64, 3, seed=1)
    t = lambda a: torch.from_numpy(a).cuda()
    tc, tj, te, m = t(cost), t(jump), t(edge), t(M0)
    torch.cuda.synchronize()
    nat.timing_reset()
    nat.timing_enable(True)
    try:
        for _ in range(3):
            st.forward_minsum(tc, tj, te, m, 3)
        torch.cuda.synchronize()
    finally:
        nat.timing_enable(False)
    got = nat.timing_read()
    assert set(got) == {"lbp_fwd_horizontal", "lbp_fwd_vertical"}, got
    for name, (ms, n) in got.items():
        assert n == 3 and ms > 0.0, (name, ms, n)
    # disabled: further launches are not booked
    st.forward_minsum(tc, tj, te, m, 3)
    torch.cuda.synchronize()
    assert nat.timing_read() == got
    nat.timing_reset()
    assert nat.timing_read() == {}
