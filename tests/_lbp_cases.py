"""Shared builders for LBP test problems (plain numpy; used by CPU and GPU tests)."""
import numpy as np

LEFT, RIGHT, UP, DOWN = 0, 1, 2, 3


def pack_stereo_jump(aff, offset=None):
    """Closed form of the K_in=3 / K_in=5 affinity packing (message_passing_op_cuda.py:35-56).

    aff [N,4,K_in,H,W] -> jump [N,4,K_in+3,H,W].
    K_in=3, aff=(a0,a1,L2): LEFT/UP [o,L2,a0,0,a1,L2], RIGHT/DOWN [o,L2,a1,0,a0,L2].
    K_in=5, aff=(L2-,L2+,L1-,L1+,L3): LEFT/UP [o,L3,L2-,L1-,0,L1+,L2+,L3],
    RIGHT/DOWN [o,L3,L2+,L1+,0,L1-,L2-,L3].
    """
    aff = np.asarray(aff, dtype=np.float64)
    N, _, Ki, H, W = aff.shape
    J = np.zeros((N, 4, Ki + 3, H, W))
    if offset is not None:
        J[:, :, 0] = offset
    if Ki == 3:
        a0, a1, L2 = aff[:, :, 0], aff[:, :, 1], aff[:, :, 2]
        for d in range(4):
            lo, hi = (a0, a1) if d in (LEFT, UP) else (a1, a0)
            J[:, d, 1] = L2[:, d]
            J[:, d, 2] = lo[:, d]
            J[:, d, 4] = hi[:, d]
            J[:, d, 5] = L2[:, d]
    elif Ki == 5:
        L2m, L2p, L1m, L1p, L3 = (aff[:, :, i] for i in range(5))
        for d in range(4):
            if d in (LEFT, UP):
                seq = (L3, L2m, L1m, None, L1p, L2p, L3)
            else:
                seq = (L3, L2p, L1p, None, L1m, L2m, L3)
            for k, v in enumerate(seq, start=1):
                if v is not None:
                    J[:, d, k] = v[:, d]
    else:
        raise ValueError(Ki)
    return J


def unpack_stereo_grad(gJ, K_in):
    """Adjoint of pack_stereo_jump (message_passing_op_cuda.py:75-91): grad wrt aff."""
    gJ = np.asarray(gJ)
    N, _, K, H, W = gJ.shape
    g = np.zeros((N, 4, K_in, H, W))
    if K_in == 3:
        for d in range(4):
            lo, hi = (0, 1) if d in (LEFT, UP) else (1, 0)
            g[:, d, lo] = gJ[:, d, 2]
            g[:, d, hi] = gJ[:, d, 4]
            g[:, d, 2] = gJ[:, d, 1] + gJ[:, d, 5]
    else:
        for d in range(4):
            if d in (LEFT, UP):
                order = (4, 0, 2, None, 3, 1, 4)
            else:
                order = (4, 1, 3, None, 2, 0, 4)
            for k, a in enumerate(order, start=1):
                if a is not None:
                    g[:, d, a] += gJ[:, d, k]
    return g


def random_stereo_problem(rng, N, H, W, C, K_in=3, ordered=True, scale=1.0, edges="random"):
    cost = rng.normal(size=(N, H, W, C)) * scale
    if K_in == 3:
        a = rng.uniform(0.05, 0.6, size=(N, 4, 2, H, W))
        L2 = rng.uniform(0.2, 1.0, size=(N, 4, 1, H, W))
        if ordered:  # strict margin: equal entries would put exact ties (kinks) in the problem
            L2 = np.maximum(L2, a.max(axis=2, keepdims=True) + 0.05)
        aff = np.concatenate([a, L2], axis=2)
    else:
        aff = rng.uniform(0.05, 1.0, size=(N, 4, 5, H, W))
        if ordered:
            aff[:, :, 0] = np.maximum(aff[:, :, 0], aff[:, :, 2] + 0.05)
            aff[:, :, 1] = np.maximum(aff[:, :, 1], aff[:, :, 3] + 0.05)
            aff[:, :, 4] = np.maximum(aff[:, :, 4], np.maximum(aff[:, :, 0], aff[:, :, 1]) + 0.05)
    jump = pack_stereo_jump(aff)
    if edges == "ones":
        edge = np.ones((N, 4, H, W))
    else:
        edge = rng.uniform(0.5, 1.5, size=(N, 4, H, W))
    return cost, aff, jump, edge


def semantic_matrix(rng, C, zero_diag=True):
    M = rng.uniform(0.0, 1.0, size=(C, C))
    if zero_diag:
        np.fill_diagonal(M, 0.0)
    return M


def pack_semantic_global(pw_mat, single_pw=False):
    """jump_in [1,4,C,C] from pw_mat [1,2,C,C] (ops/lbp_semantic_pw/message_passing_op_cuda.py:27-40):
    LEFT=M_h, RIGHT=M_h^T, UP=M_v, DOWN=M_v^T; single_pw: all four = M_h."""
    pw_mat = np.asarray(pw_mat)
    if single_pw:
        return np.stack([pw_mat[0, 0]] * 4)[None]
    Mh, Mv = pw_mat[0, 0], pw_mat[0, 1]
    return np.stack([Mh, Mh.T, Mv, Mv.T])[None]
