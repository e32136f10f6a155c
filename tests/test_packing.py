"""CPU checks of the Python glue around the ops (packing, shifts, adjoint unpacking)
against closed forms and adjoint identities."""
import numpy as np
import torch

from _lbp_cases import pack_stereo_jump, unpack_stereo_grad, pack_semantic_global


def test_stereo_packing_matches_closed_form():
    from bplayers_amd.ops.lbp_stereo.message_passing_op_cuda import pack_affinities, unpack_affinity_grad
    rng = np.random.default_rng(0)
    for Ki in (3, 5):
        aff = rng.normal(size=(2, 4, Ki, 3, 5))
        off = rng.normal(size=(2, 4, 3, 5))
        J = pack_affinities(torch.tensor(aff), torch.tensor(off)).numpy()
        np.testing.assert_allclose(J, pack_stereo_jump(aff, off))
        gJ = rng.normal(size=J.shape)
        ga, go = unpack_affinity_grad(torch.tensor(gJ))
        np.testing.assert_allclose(ga.numpy(), unpack_stereo_grad(gJ, Ki))
        # adjoint identity <pack(a), gJ> = <a, unpack(gJ)> (+ offset part)
        lhs = (J * gJ).sum()
        rhs = (aff * ga.numpy()).sum() + (off * go.numpy()).sum()
        np.testing.assert_allclose(lhs, rhs)


def test_stereo_packing_gives_the_p1p2_pairwise():
    """f(c,l) = J[clamp(K/2 + c - l, 1, K-1)] reproduces construct_pw_tensor's P1/P2 matrix."""
    from bplayers_amd.ops.lbp_stereo.message_passing_op_cuda import pack_affinities, construct_pw_tensor
    L1, L2, C = 0.1, 0.5, 7
    aff = torch.zeros((1, 4, 3, 1, 1))
    aff[:, :, :2] = L1
    aff[:, :, 2] = L2
    J = pack_affinities(aff, torch.zeros((1, 4, 1, 1)))[0, 0, :, 0, 0]
    f = np.array([[float(J[min(max(3 + c - l, 1), 5)]) for l in range(C)] for c in range(C)])
    np.testing.assert_allclose(f, construct_pw_tensor(L1, L2, C)[0, 0].numpy(), atol=1e-7)


def test_semantic_global_packing_and_adjoint():
    from bplayers_amd.ops.lbp_semantic_pw.message_passing_op_cuda import pack_pairwise, unpack_pairwise_grad
    rng = np.random.default_rng(1)
    pw = rng.normal(size=(1, 2, 5, 5))
    for single in (False, True):
        J = pack_pairwise(torch.tensor(pw), single).numpy()
        np.testing.assert_allclose(J, pack_semantic_global(pw, single))
        gJ = rng.normal(size=J.shape)
        g = unpack_pairwise_grad(torch.tensor(gJ), single).numpy()
        src = pw if not single else pw[:, :1]
        np.testing.assert_allclose((J * gJ).sum(), (src * g).sum())


def test_semantic_pixel_shift_and_adjoint():
    from bplayers_amd.ops.lbp_semantic_pw_pixel.message_passing_op_pw_pixel import shift_pairwise, fold_pairwise_grad
    rng = np.random.default_rng(2)
    j = rng.normal(size=(2, 3, 4, 5, 5))
    s = shift_pairwise(torch.tensor(j)).numpy()
    np.testing.assert_allclose(s[0], j[0])
    np.testing.assert_allclose(s[2], j[1])
    np.testing.assert_allclose(s[1][:, 1:], np.swapaxes(j[0][:, :-1], -1, -2))
    assert np.all(s[1][:, 0] == 0)
    np.testing.assert_allclose(s[3][1:], np.swapaxes(j[1][:-1], -1, -2))
    assert np.all(s[3][0] == 0)
    g = rng.normal(size=s.shape)
    f = fold_pairwise_grad(torch.tensor(g)).numpy()
    np.testing.assert_allclose((s * g).sum(), (j * f).sum())


def test_bp_adjust_inputs_closed_forms():
    from bplayers_amd.ops.lbp_stereo.bp_op_cuda import BP
    rng = np.random.default_rng(3)
    bp = BP.__new__(BP)
    w = torch.tensor(rng.uniform(size=(1, 2, 3, 4)), dtype=torch.float32)
    wi = BP.adjust_input_weights(bp, w, 0).numpy()
    np.testing.assert_allclose(wi[:, 0], w[:, 0])
    np.testing.assert_allclose(wi[:, 1, :, 1:], w[:, 0, :, :-1])
    assert np.all(wi[:, 1, :, 0] == 0)
    np.testing.assert_allclose(wi[:, 3, 1:], w[:, 1, :-1])
    a = torch.tensor(rng.uniform(size=(1, 2, 5, 3, 4)), dtype=torch.float32)
    s = BP.adjust_input_affinities(bp, a).numpy()
    assert np.all(s[:, 0, 0] >= s[:, 0, 2]) and np.all(s[:, 0, 4] >= s[:, 0, 0]) and np.all(s[:, 0, 4] >= s[:, 0, 1])
    np.testing.assert_allclose(s[:, 1, :, :, 1:], s[:, 0, :, :, :-1])
    o = torch.tensor(rng.normal(size=(1, 2, 3, 4)), dtype=torch.float32)
    so = BP.adjust_input_offsets(bp, o).numpy()
    np.testing.assert_allclose(so[:, 1, :, 1:], -o.numpy()[:, 0, :, :-1])
    np.testing.assert_allclose(so[:, 3, 1:], -o.numpy()[:, 1, :-1])


def test_sub_expectation_closed_form():
    from bplayers_amd.ops.lbp_stereo.inference_op import Inference
    b = torch.zeros((1, 1, 1, 10))
    b[0, 0, 0, 4] = 0.5
    b[0, 0, 0, 5] = 0.3
    b[0, 0, 0, 2] = 0.2
    b[0, 0, 0, 9] = 0.9  # outside the +-3 window of... the argmax is 9 here
    r = Inference.compute_sub_expectation(b)
    win = [6, 7, 8, 9, 9, 9, 9]
    vals = b[0, 0, 0, win]
    np.testing.assert_allclose(float(r), float((vals * torch.tensor(win, dtype=torch.float32)).sum() / vals.sum()),
                               rtol=1e-6)


def test_pairwise_tables_known_answers():
    """Closed-form P1/P2 tables and the jump-vector lookup on hand-computed values."""
    from bplayers_amd.ops._pairwise import jump_index, jump_value, smoothness_prob, truncated_linear
    from bplayers_amd.ops.lbp_stereo.bp_op_cuda import BP
    E = truncated_linear(4, 0.1, 0.5)
    np.testing.assert_allclose(E, [[0, .1, .5, .5], [.1, 0, .1, .5], [.5, .1, 0, .1], [.5, .5, .1, 0]])
    P = smoothness_prob(3, 0.6, 0.15)
    np.testing.assert_allclose(P, [[.6, .15, .25], [.15, .6, .15], [.25, .15, .6]])
    # a zero weight is replaced by 1 - L1 - L2 like every other zero entry
    np.testing.assert_allclose(smoothness_prob(2, 0.0, 0.3), [[.7, .3], [.3, .7]])
    assert list(jump_index(np.arange(-4, 5), 6)) == [1, 1, 1, 2, 3, 4, 5, 5, 5]
    J = np.array([0.5, 2.0, 1.0, 0.0, 1.5, 3.0])  # offset 0.5
    assert jump_value(1, 1, J) == 0.75            # between J[3] = 0 and J[4] = 1.5
    assert jump_value(3, 1, J) == 3.0             # 2.5: J[5] and the clamped J[5]
    assert jump_value(0, 2, J) == 0.5 * 2.0 + 0.5 * 1.0  # -1.5: J[1] and J[2]
    J0 = J.copy()
    J0[0] = 0.0
    assert jump_value(2, 4, J0) == J0[1] and jump_value(4, 3, J0) == J0[4]  # integer differences
    assert BP.f_func is not None and BP.construct_pw_energy_weights(3, 0.1, 0.2)[0, 2] == 0.2
