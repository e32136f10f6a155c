"""The C++/OpenMP restatement (oracle/c) agrees with the numpy oracle: bitwise-identical
argmins and float64 values to rounding, forward and exact backward."""
import numpy as np
import pytest

from oracle import cpu
from oracle import lbp as O
from _lbp_cases import random_stereo_problem


@pytest.mark.parametrize("C,delta,ordered,K_in", [(9, 3, False, 3), (16, 3, True, 5), (6, 6, True, 3)])
def test_cpu_restatement_matches_numpy_oracle(C, delta, ordered, K_in):
    rng = np.random.default_rng(C)
    N, H, W = 2, 4, 6
    cost, aff, jump, edge = random_stereo_problem(rng, N, H, W, C, K_in=K_in, ordered=ordered)
    M0 = rng.normal(size=(N, 4, H, W, C)) * 0.3
    Mc, ac, fc = cpu.lbp_forward(cost, jump, edge, M0, delta, dtype=np.float64)
    Mo, ao, so, fo, lso = O.lbp_forward(cost, O.StereoPairwise(jump), edge, M0, delta)
    np.testing.assert_allclose(Mc, Mo, atol=1e-12)
    lab_o = np.where(fo, lso[..., None], ao)
    np.testing.assert_array_equal(ac, np.where(fo, lab_o, ao))
    np.testing.assert_array_equal(fc.astype(bool), fo)
    G = rng.normal(size=Mo.shape)
    gc, gj, ge, gm = cpu.lbp_backward_exact(jump, edge, G, Mc, ac, fc, delta, dtype=np.float64)
    oc, op, oe, om = O.lbp_backward_exact(cost, O.StereoPairwise(jump), edge, Mo, ao, fo, G)
    for a, b in ((gc, oc), (gj, op), (ge, oe), (gm, om)):
        np.testing.assert_allclose(a, b, atol=1e-10)


def test_cpu_sad_matches_oracle():
    from oracle import sad as S
    rng = np.random.default_rng(0)
    f0 = rng.normal(size=(1, 5, 3, 20)).astype(np.float32)
    f1 = rng.normal(size=(1, 5, 3, 20)).astype(np.float32)
    np.testing.assert_allclose(cpu.sad_forward(f0, f1, 0, 7), S.stereo_sad_forward(f0, f1, 0, 7), rtol=1e-5)
