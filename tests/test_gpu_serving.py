"""GPU tests of the continuous-batching stereo server (bplayers_amd.serving)."""
import numpy as np
import pytest
import torch

import _parity as P
from oracle import lbp as O
from oracle import sad as S
from _lbp_cases import pack_stereo_jump

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")
SHAPES = [(24, 40, 16), (32, 48, 24), (20, 56, 32)]
C_FEAT = 8


def _pool(seed=0, per_shape=2):
    from bplayers_amd import workloads
    pool = {}
    for i, (H, W, D) in enumerate(SHAPES):
        pool[(H, W, D)] = []
        for j in range(per_shape):
            f0, f1, _ = workloads.stereo_pair(seed + 10 * i + j, C_FEAT, H, W, 0.75 * D)
            pool[(H, W, D)].append((torch.from_numpy(f0).to(DEV), torch.from_numpy(f1).to(DEV)))
    return pool


def test_result_is_the_same_alone_and_in_the_mix():
    from bplayers_amd.serving import StereoServer, make_requests
    pool = _pool()
    reqs_a = make_requests(pool, 24, rate_hz=1e5, seed=1)
    reqs_b = make_requests(pool, 24, rate_hz=1e5, seed=1)
    seq = StereoServer(DEV, C=C_FEAT, hbm_budget=1 << 30, sequential=True)
    st_a = seq.serve(reqs_a, keep_results=True)
    mix = StereoServer(DEV, C=C_FEAT, hbm_budget=1 << 30, max_batch=4, max_slots_per_class=2)
    st_b = mix.serve(reqs_b, keep_results=True)
    assert st_a["frames"] == st_b["frames"] == 24
    for ra, rb in zip(sorted(reqs_a, key=lambda r: r.rid), sorted(reqs_b, key=lambda r: r.rid)):
        assert torch.equal(ra.result, rb.result), f"request {ra.rid} differs alone vs in the mix"
    print("sequential", st_a, "\nmixed", st_b)


def test_served_disparity_matches_oracle_pipeline():
    from bplayers_amd.serving import StereoServer, Request
    pool = _pool(per_shape=1)
    shape = SHAPES[1]
    f0, f1 = pool[shape][0]
    srv = StereoServer(DEV, C=C_FEAT, hbm_budget=1 << 30, sequential=True)
    r = Request(0, shape, f0, f1, 0.0)
    srv.serve([r], keep_results=True)
    H, W, D = shape
    sad = S.stereo_sad_forward(f0.cpu().numpy(), f1.cpu().numpy(), 0, D - 1)
    z = -sad
    prob = np.exp(z - z.max(-1, keepdims=True))
    prob /= prob.sum(-1, keepdims=True)
    cost = -prob
    aff = np.zeros((1, 4, 3, H, W))
    aff[:, :, :2] = 0.1
    aff[:, :, 2] = 0.5
    M, *_ = O.lbp_forward(cost, O.StereoPairwise(pack_stereo_jump(aff)), np.ones((1, 4, H, W)),
                          np.zeros((1, 4, H, W, D)), 3)
    E = (cost + M.sum(axis=1))[0]
    lab = E.argmin(-1)
    # a one-iteration solve's energies carry the rounding of the SAD sums through the softmax
    # and of four messages: bounded by 1e-5 of their magnitudes (the per-vector message bar)
    dev = 1e-5 * (np.abs(cost).max() + 4 * np.abs(M).max())
    clear = P.winner_clear(E, dev)
    got = r.result.cpu().numpy()
    print(f"    disparities: {1 - clear.mean():.2e} of pixels within {2 * dev:.3g} of a tie", flush=True)
    assert clear.mean() > 0.99
    np.testing.assert_array_equal(got[clear], lab[clear])


def test_admission_rejects_requests_above_the_budget():
    from bplayers_amd.serving import StereoServer, Request, AdmissionError
    pool = _pool(per_shape=1)
    shape = SHAPES[0]
    f0, f1 = pool[shape][0]
    srv = StereoServer(DEV, C=C_FEAT, hbm_budget=1 << 10)
    with pytest.raises(AdmissionError):
        srv.serve([Request(0, shape, f0, f1, 0.0)])


def test_budget_limits_resident_slots_and_still_serves_everything():
    from bplayers_amd.serving import StereoServer, make_requests, slot_bytes
    pool = _pool()
    budget = max(slot_bytes(s, C_FEAT, 2) for s in SHAPES) + 1
    srv = StereoServer(DEV, C=C_FEAT, hbm_budget=budget, max_batch=2, max_slots_per_class=2)
    stats = srv.serve(make_requests(pool, 16, rate_hz=1e5, seed=3))
    assert stats["frames"] == 16
    assert srv.resident <= budget


def test_inflight_cap_keeps_results_and_bounds_concurrency():
    """max_inflight=1 still batches requests of one class but keeps one batch on the device at
    a time; every result equals the sequential server's."""
    from bplayers_amd import serving
    from bplayers_amd.serving import StereoServer, make_requests
    pool = _pool(seed=3)
    reqs_a = make_requests(pool, 20, rate_hz=1e9, seed=4)
    reqs_b = make_requests(pool, 20, rate_hz=1e9, seed=4)
    StereoServer(DEV, C=C_FEAT, hbm_budget=1 << 30, sequential=True).serve(reqs_a, keep_results=True)
    srv = StereoServer(DEV, C=C_FEAT, hbm_budget=1 << 30, max_batch=4, max_slots_per_class=2, max_inflight=1)
    peak = [0]
    launch = serving.Slot.launch

    def counting_launch(slot, reqs):
        busy = sum(1 for lst in srv.slots.values() for s in lst if s.reqs is not None)
        peak[0] = max(peak[0], busy + 1)
        return launch(slot, reqs)
    serving.Slot.launch = counting_launch
    try:
        st = srv.serve(reqs_b, keep_results=True)
    finally:
        serving.Slot.launch = launch
    assert st["frames"] == 20 and peak[0] == 1
    assert any(len(s.reqs or []) for lst in srv.slots.values() for s in lst) is False
    for ra, rb in zip(sorted(reqs_a, key=lambda r: r.rid), sorted(reqs_b, key=lambda r: r.rid)):
        assert torch.equal(ra.result, rb.result), f"request {ra.rid} differs"

