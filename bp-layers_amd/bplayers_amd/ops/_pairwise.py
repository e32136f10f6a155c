"""Closed forms of the pairwise terms used around the LBP ops.

The stereo op reads its pairwise term as a per-pixel vector J over label differences:
f(c, l) = J[clamp(K//2 + c - l, 1, K-1)], entry 0 being the window offset (reference
ops/lbp_stereo/src/util.cuh:103-119).  A truncated-linear (P1/P2) term therefore has
J = [offset, P2, ..., P1, 0, P1, ..., P2].  This module states these tables once:

* ``label_distance`` / ``truncated_linear`` / ``smoothness_prob``: dense K x K matrices of
  the global P1/P2 model (reference message_passing_op_cuda.py:9-20 and the two static
  weight constructors of bp_op_cuda.py:53-78), from |i - j|;
* ``jump_index`` / ``jump_value``: the lookup of a jump vector at a (possibly fractional)
  label difference, with linear interpolation between the two neighbouring entries
  (bp_op_cuda.py:18-51);
* ``stereo_jump_layout``: where each affinity of the K_in = 3 / 5 parametrisations lands in
  the packed jump vector (message_passing_op_cuda.py:35-56), used by pack/unpack.
"""
import numpy as np
import torch

# source codes of a packed jump slot besides an affinity index
SLOT_OFFSET = -1
SLOT_ZERO = -2


def label_distance(K):
    """|i - j| for labels i, j in [0, K) as a K x K integer matrix."""
    r = np.arange(K)
    return np.abs(r[:, None] - r[None, :])


def truncated_linear(K, L1, L2):
    """K x K energies of the P1/P2 model: 0 on the diagonal, L1 one label apart, L2 further."""
    d = label_distance(K)
    return np.where(d == 0, 0.0, np.where(d == 1, L1, L2)).astype(np.float64)


def smoothness_prob(K, L1, L2):
    """K x K probability-style weights: L1 on the diagonal, L2 one label apart and
    1 - L1 - L2 elsewhere.  Entries that come out exactly 0 (L1 or L2 equal to 0) also take
    1 - L1 - L2, as the reference's constructor replaces every zero."""
    d = label_distance(K)
    m = np.where(d == 0, L1, np.where(d == 1, L2, 0.0)).astype(np.float64)
    m[m == 0] = 1.0 - L1 - L2
    return m


def jump_index(pos, K):
    """Entry of a length-K jump vector addressed by label difference ``pos`` (integer,
    scalar or array): K//2 + pos clamped to [1, K-1] (entry 0 holds the offset)."""
    return np.clip(K // 2 + np.asarray(pos), 1, K - 1)


def jump_value(t, s, J):
    """Pairwise value of receiver label t and sender label s for the jump vector J
    (1-D, J[0] = window offset).  The difference t - s + J[0] is generally fractional; the
    value is the linear interpolation of the entries at its floor and ceiling (an integer
    difference reads one entry)."""
    J = np.asarray(J, dtype=np.float64)
    K = J.shape[0]
    p = np.asarray(t, dtype=np.float64) - np.asarray(s, dtype=np.float64) + J[0]
    lo = np.floor(p)
    frac = p - lo
    v_lo = J[jump_index(lo.astype(np.int64), K)]
    v_hi = J[jump_index(lo.astype(np.int64) + 1, K)]
    out = np.where(frac == 0, v_lo, (1.0 - frac) * v_lo + frac * v_hi)
    return out if out.ndim else float(out)


def stereo_jump_layout(K_in):
    """Source of every slot of the packed jump vector (length K_in + 3) for the two
    direction parities: a tuple (even, odd) of lists whose entries are an affinity index,
    SLOT_OFFSET or SLOT_ZERO.

    Directions LEFT and UP (even) take the affinities in their stored order around the
    centre, RIGHT and DOWN (odd) swap each (minus, plus) pair, so that an odd direction's
    term is the transpose of the even one's; both ends hold the last affinity (the
    truncation value) and the centre is 0."""
    if K_in not in (3, 5):
        raise ValueError(f"affinity vectors have 3 or 5 entries, got {K_in}")
    K = K_in + 3
    even = [SLOT_ZERO] * K
    even[0] = SLOT_OFFSET
    even[1] = even[K - 1] = K_in - 1
    odd = list(even)
    for pair, slot in enumerate(range(2, K // 2)):
        minus, plus = 2 * pair, 2 * pair + 1
        even[slot], even[K - slot] = minus, plus
        odd[slot], odd[K - slot] = plus, minus
    return even, odd


def _gather_table(K_in):
    """Index tables into the extended affinity axis [affinities..., offset, zero]."""
    ext = {SLOT_OFFSET: K_in, SLOT_ZERO: K_in + 1}
    return [torch.tensor([ext.get(s, s) for s in tab], dtype=torch.long) for tab in stereo_jump_layout(K_in)]


def pack_stereo_jump(affinities, offset):
    """[N,4,K_in,H,W] affinities + [N,4,H,W] offset -> [N,4,K_in+3,H,W] jump vectors."""
    N, D, K_in, H, W = affinities.shape
    zero = torch.zeros_like(offset)
    ext = torch.cat([affinities, offset.to(affinities)[:, :, None], zero.to(affinities)[:, :, None]], dim=2)
    even, odd = (t.to(affinities.device) for t in _gather_table(K_in))
    J = torch.empty((N, D, K_in + 3, H, W), dtype=affinities.dtype, device=affinities.device)
    J[:, 0::2] = ext[:, 0::2].index_select(2, even)
    J[:, 1::2] = ext[:, 1::2].index_select(2, odd)
    return J.contiguous()


def unpack_stereo_jump_grad(grad_jump):
    """Adjoint of pack_stereo_jump: (grad_affinities [N,4,K_in,H,W], grad_offset [N,4,H,W])."""
    N, D, K, H, W = grad_jump.shape
    K_in = K - 3
    even, odd = (t.to(grad_jump.device) for t in _gather_table(K_in))
    g = torch.zeros((N, D, K_in + 2, H, W), dtype=grad_jump.dtype, device=grad_jump.device)
    g[:, 0::2] = g[:, 0::2].index_add(2, even, grad_jump[:, 0::2])
    g[:, 1::2] = g[:, 1::2].index_add(2, odd, grad_jump[:, 1::2])
    return g[:, :, :K_in].contiguous(), g[:, :, K_in].contiguous()
