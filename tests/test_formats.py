"""File formats (CPU): .flo, PFM, and the reference's parameter checkpoints."""
import glob
import os

import numpy as np
import pytest
import torch

from bplayers_amd import formats

REF_DATA = "/root/reference/data"


def test_flo_round_trip_and_layout(tmp_path):
    rng = np.random.default_rng(0)
    flow = rng.normal(size=(5, 7, 2)).astype(np.float32)
    p = tmp_path / "a.flo"
    formats.write_flo(p, flow)
    raw = p.read_bytes()
    assert raw[:4] == b"PIEH" and np.frombuffer(raw[:4], "<f4")[0] == np.float32(202021.25)
    assert tuple(np.frombuffer(raw[4:12], "<i4")) == (7, 5)
    assert np.frombuffer(raw[12:28], "<f4").tolist() == flow[0, 0].tolist() + flow[0, 1].tolist()
    np.testing.assert_array_equal(formats.read_flo(p), flow)
    p.write_bytes(b"XXXX" + raw[4:])
    with pytest.raises(ValueError, match="PIEH"):
        formats.read_flo(p)


@pytest.mark.skipif(not os.path.exists(os.path.join(REF_DATA, "frame_0019.flo")), reason="reference data absent")
def test_reads_the_reference_sample_flow():
    path = os.path.join(REF_DATA, "frame_0019.flo")
    flow = formats.read_flo(path)
    w, h = np.fromfile(path, "<i4", count=3)[1:]
    assert flow.shape == (h, w, 2) and os.path.getsize(path) == 12 + h * w * 8
    assert np.isfinite(flow).all()


def test_pfm_layout_and_round_trip(tmp_path):
    img = np.arange(6, dtype=np.float32).reshape(2, 3)  # row 0 = top
    p = tmp_path / "d.pfm"
    formats.write_pfm(p, img)
    raw = p.read_bytes()
    head = b"Pf\n3 2\n-1.0\n"
    assert raw[:len(head)] == head
    # the format stores the bottom row first
    assert np.frombuffer(raw[len(head):], "<f4").tolist() == [3, 4, 5, 0, 1, 2]
    np.testing.assert_array_equal(formats.read_pfm(p), img)


def test_legacy_bp_checkpoint_is_adjusted(tmp_path):
    """An older BP checkpoint (L1/L2 without the underscore, no softmin temperature) loads
    through the reference's adjustment hook."""
    from bplayers_amd.ops.lbp_stereo.bp_op_cuda import BP
    bp = BP("cpu", max_iter=1, num_labels=8)
    p = tmp_path / "crf.cpt"
    torch.save({"message_passing.L1": torch.tensor(0.3), "message_passing.L2": torch.tensor(1.7)}, p)
    formats.load_checkpoint(bp, p, "cpu")
    assert float(bp.message_passing._L1.detach()) == pytest.approx(0.3)
    assert float(bp.message_passing._L2.detach()) == pytest.approx(1.7)
    assert float(bp.message_passing.softmin.T.detach()) == 1.0


CRF = sorted(glob.glob(os.path.join(REF_DATA, "params", "*", "*", "crf0_lvl*_best.cpt")))
MATCH = sorted(glob.glob(os.path.join(REF_DATA, "params", "*", "*", "matching_lvl*_best.cpt")))


@pytest.mark.skipif(not CRF, reason="reference checkpoints absent")
@pytest.mark.parametrize("path", CRF)
def test_loads_the_reference_bp_checkpoints(path):
    from bplayers_amd.ops.lbp_stereo.bp_op_cuda import BP
    bp = BP("cpu", max_iter=1, num_labels=8)
    formats.load_checkpoint(bp, path, "cpu")
    sd = torch.load(path, map_location="cpu", weights_only=True)
    for k, v in sd.items():
        assert torch.equal(bp.state_dict()[k], v), k


@pytest.mark.skipif(not MATCH, reason="reference checkpoints absent")
@pytest.mark.parametrize("path", MATCH)
def test_loads_the_reference_matching_checkpoints(path):
    from bplayers_amd.matching import StereoMatchingSad
    m = StereoMatchingSad(0, 31, device="cpu")
    formats.load_checkpoint(m, path, "cpu")
    sd = torch.load(path, map_location="cpu", weights_only=True)
    assert torch.equal(m._softmax.T.detach(), sd["_softmax.T"])


FLOW_MATCH = [p for p in MATCH if os.sep + "flow" + os.sep in p]


@pytest.mark.skipif(not FLOW_MATCH, reason="reference checkpoints absent")
@pytest.mark.parametrize("path", FLOW_MATCH)
def test_loads_the_reference_flow_matching_checkpoints(path):
    from bplayers_amd.flow_matching import FlowMatchingSad
    m = FlowMatchingSad(24, device="cpu")
    formats.load_checkpoint(m, path, "cpu")
    sd = torch.load(path, map_location="cpu", weights_only=True)
    assert torch.equal(m._softmax.T.detach(), sd["_softmax.T"])

