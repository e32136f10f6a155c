"""Pins the FP64 SAD oracles (oracle/sad.py): scalar brute-force loops written from the
kernel text, and finite differences of the forward for the backward."""
import numpy as np
import pytest

from oracle import sad as S


def brute_stereo(f0, f1, min_disp, max_disp):
    N, C, H, W = f0.shape
    D = max_disp - min_disp + 1
    out = np.full((N, H, W, D), 40.0)
    for n in range(N):
        for y in range(H):
            for x in range(W):
                for i in range(D):
                    d = i + min_disp
                    if 0 <= x - d < W:
                        out[n, y, x, i] = sum(abs(f0[n, c, y, x] - f1[n, c, y, x - d]) for c in range(C))
    return out


@pytest.mark.parametrize("mn,mx", [(0, 5), (-2, 3), (2, 9)])
def test_stereo_forward_brute(mn, mx):
    rng = np.random.default_rng(1)
    f0 = rng.normal(size=(2, 3, 2, 7))
    f1 = rng.normal(size=(2, 3, 2, 7))
    np.testing.assert_allclose(S.stereo_sad_forward(f0, f1, mn, mx), brute_stereo(f0, f1, mn, mx), atol=1e-12)


def test_stereo_backward_matches_fd():
    rng = np.random.default_rng(2)
    f0 = rng.normal(size=(1, 3, 2, 8))
    f1 = rng.normal(size=(1, 3, 2, 8))
    mn, mx = -1, 4
    G = rng.normal(size=(1, 2, 8, mx - mn + 1))
    df0, df1 = S.stereo_sad_backward(f0, f1, mn, mx, G)
    eps = 1e-7
    for arr, g in ((f0, df0), (f1, df1)):
        flat = arr.reshape(-1)
        for i in range(flat.size):
            o = flat[i]
            flat[i] = o + eps
            lp = (S.stereo_sad_forward(f0, f1, mn, mx) * G).sum()
            flat[i] = o - eps
            lm = (S.stereo_sad_forward(f0, f1, mn, mx) * G).sum()
            flat[i] = o
            assert abs((lp - lm) / (2 * eps) - g.reshape(-1)[i]) < 1e-6


def test_stereo_num_disp_and_step():
    assert S.stereo_num_disp(0, 63) == 64
    assert S.stereo_num_disp(0, 191) == 192
    assert S.stereo_num_disp(0, 9, 0.5) == 20
    assert [S.stereo_disp(i, 0, 0.5) for i in range(5)] == [0, 0, 1, 1, 2]


def brute_flow(f0, f1, sws, ou=0, ov=0, bu=0, bv=0):
    N, C, H, W = f0.shape
    K = sws + 1
    h = sws // 2
    cvu = np.zeros((N, H, W, K)); cvv = np.zeros((N, H, W, K))
    us = np.zeros((N, H, W, K)); vs = np.zeros((N, H, W, K))
    for n in range(N):
        for y in range(H):
            for x in range(W):
                Sm = np.full((K, K), 40.0)  # [v_idx, u_idx]
                for ui in range(K):  # one thread per u_idx in [0, K)
                    u = ui - h
                    if not (0 <= x + ou + u < W):
                        continue
                    for v in range(-h, h + 1):  # v over [-h, h]: v_idx = K-1 stays 40 for odd sws
                        vi = v + h
                        if not (0 <= y + ov + v < H):
                            continue
                        Sm[vi, ui] = sum(abs(f0[n, c, y, x] - f1[n, c, y + ov + v, x + ou + u]) for c in range(C))
                for ui in range(K):
                    best, arg = 9999999.0, 0
                    for vi in range(K):
                        if Sm[vi, ui] < best:
                            best, arg = Sm[vi, ui], vi
                    cvu[n, y, x, ui] = best
                    us[n, y, x, ui] = arg + bv * sws
                for vi in range(K):
                    best, arg = 9999999.0, 0
                    for ui in range(K):
                        if Sm[vi, ui] < best:
                            best, arg = Sm[vi, ui], ui
                    cvv[n, y, x, vi] = best
                    vs[n, y, x, vi] = arg + bu * sws
    return cvu, cvv, us, vs


@pytest.mark.parametrize("sws,ou,ov,bu,bv", [(4, 0, 0, 0, 0), (2, 1, -1, 1, 0), (6, 0, 0, 0, 1), (3, 0, 0, 0, 0),
                                            (5, -1, 2, 1, 1)])
def test_flow_forward_brute(sws, ou, ov, bu, bv):
    rng = np.random.default_rng(3)
    f0 = rng.normal(size=(1, 2, 4, 5))
    f1 = rng.normal(size=(1, 2, 4, 5))
    got = S.flow_sad_forward(f0, f1, sws, ou, ov, bu, bv)
    exp = brute_flow(f0, f1, sws, ou, ov, bu, bv)
    for g, e in zip(got, exp):
        np.testing.assert_allclose(g, e, atol=1e-12)


def test_flow_forward_ties_take_first_index():
    f0 = np.zeros((1, 1, 3, 3))
    f1 = np.zeros((1, 1, 3, 3))
    cvu, cvv, us, vs = S.flow_sad_forward(f0, f1, 2)
    # every in-range S is 0: first in-range v / u wins
    assert us[0, 0, 0, 1] == 1 and vs[0, 0, 0, 1] == 1  # v=-1 / u=-1 are outside at (0,0)
    assert us[0, 1, 1, 0] == 0 and vs[0, 1, 1, 0] == 0


@pytest.mark.parametrize("sws", [4, 3])
def test_flow_backward_matches_fd(sws):
    """The backward is the derivative of the min-projected volumes with the argmins fixed
    (for odd sws the last label carries no gradient, so its incoming gradient is zeroed)."""
    rng = np.random.default_rng(4)
    f0 = rng.normal(size=(1, 2, 5, 6))
    f1 = rng.normal(size=(1, 2, 5, 6))
    cvu, cvv, us, vs = S.flow_sad_forward(f0, f1, sws)
    Gu = rng.normal(size=cvu.shape)
    Gv = rng.normal(size=cvv.shape)
    Gu[..., 2 * (sws // 2) + 1:] = 0
    Gv[..., 2 * (sws // 2) + 1:] = 0
    df0, df1 = S.flow_sad_backward(f0, f1, sws, Gu, Gv, us, vs)
    eps = 1e-7

    def L():
        a, b, _, _ = S.flow_sad_forward(f0, f1, sws)
        return (a * Gu).sum() + (b * Gv).sum()
    for arr, g in ((f0, df0), (f1, df1)):
        flat = arr.reshape(-1)
        for i in range(flat.size):
            o = flat[i]
            flat[i] = o + eps
            lp = L()
            flat[i] = o - eps
            lm = L()
            flat[i] = o
            assert abs((lp - lm) / (2 * eps) - g.reshape(-1)[i]) < 1e-6, i


def brute_flow_backward(f0, f1, sws, gu, gv, us, vs):
    """Literal restatement of flow_mp_sad_kernel.cu:126-200: per (c, y, x), u over [-h, h]
    with v = u_star - h, then v over [-h, h] with u = v_star - h; df1 by scattered adds."""
    N, C, H, W = f0.shape
    h = sws // 2
    df0 = np.zeros_like(f0)
    df1 = np.zeros_like(f1)
    for n in range(N):
        for c in range(C):
            for y in range(H):
                for x in range(W):
                    acc = 0.0
                    for u in range(-h, h + 1):
                        v = int(us[n, y, x, u + h]) - h
                        if 0 <= x + u < W and 0 <= y + v < H:
                            d = f0[n, c, y, x] - f1[n, c, y + v, x + u]
                            if abs(d) > 1e-15:
                                upd = np.sign(d) * gu[n, y, x, u + h]
                                acc += upd
                                df1[n, c, y + v, x + u] -= upd
                    for v in range(-h, h + 1):
                        u = int(vs[n, y, x, v + h]) - h
                        if 0 <= x + u < W and 0 <= y + v < H:
                            d = f0[n, c, y, x] - f1[n, c, y + v, x + u]
                            if abs(d) > 1e-15:
                                upd = np.sign(d) * gv[n, y, x, v + h]
                                acc += upd
                                df1[n, c, y + v, x + u] -= upd
                    df0[n, c, y, x] = acc
    return df0, df1


@pytest.mark.parametrize("sws", [4, 5])
def test_flow_backward_brute_any_argmins(sws):
    """Random stored argmins, including ones outside the window (block offsets, negatives)
    and the last label of odd windows: the oracle equals the literal loops."""
    rng = np.random.default_rng(sws)
    N, C, H, W = 2, 2, 6, 7
    K = sws + 1
    f0 = rng.normal(size=(N, C, H, W))
    f1 = rng.normal(size=(N, C, H, W))
    us = rng.integers(-3, 2 * K + 2, size=(N, H, W, K)).astype(np.float64)
    vs = rng.integers(-3, 2 * K + 2, size=(N, H, W, K)).astype(np.float64)
    gu = rng.normal(size=(N, H, W, K))
    gv = rng.normal(size=(N, H, W, K))
    got = S.flow_sad_backward(f0, f1, sws, gu, gv, us, vs)
    exp = brute_flow_backward(f0, f1, sws, gu, gv, us, vs)
    for g, e in zip(got, exp):
        np.testing.assert_allclose(g, e, atol=1e-12)


def test_flow_backward_odd_window_last_label_has_no_gradient():
    rng = np.random.default_rng(11)
    f0 = rng.normal(size=(1, 2, 5, 6))
    f1 = rng.normal(size=(1, 2, 5, 6))
    sws = 5
    cvu, cvv, us, vs = S.flow_sad_forward(f0, f1, sws)
    Gu = np.zeros_like(cvu)
    Gv = np.zeros_like(cvv)
    Gu[..., sws] = 1.0
    Gv[..., sws] = 1.0
    df0, df1 = S.flow_sad_backward(f0, f1, sws, Gu, Gv, us, vs)
    assert not df0.any() and not df1.any()
    # the odd window's last v row is the fill value everywhere
    S5 = S.flow_sad_forward(f0, f1, sws)
    assert (S5[1][..., sws] == S.FILL).all()
