"""Multi-scale optical-flow BP around the ops (reference flow.py:80-115 and 169-182): per
pyramid level, the BP layers run on the level's two probability volumes ([N, 2, H, W, K]: u
then v, v starting from u's messages — the 5-D path of the stereo BP layer), coarse to fine;
the beliefs of each flow component are upsampled trilinearly to (2H, 2W, 2K-1) and half of
them added to the next finer level's probabilities before its BP.

Behaviour kept from the reference: the addition is out of place and happens with or without
gradients; each level's flow is its winner-take-all label minus sws // 2; the pyramids are
returned finest level first; with ``unpad`` = (l, r, t, b) the finest level's beliefs and
flow are cropped.  The reference indexes the affinities of BP layer ``idx`` as
``affinities[:, idx]`` and so needs an affinity pyramid; here ``affinity_pyramid=None``
leaves each BP layer its own P1/P2."""
import torch
import torch.nn.functional as F


def optimize_crf(crf_layers, prob_vol, weights, affinities, offsets):
    """Runs the BP layers of one level in turn (flow.py:169-182); ``affinities`` is
    [1, layers, 2, 5, H, W] or None."""
    if not crf_layers:
        return None
    flow = affinities_shift = offsets_shift = None
    for idx, crf in enumerate(crf_layers):
        prob_vol = prob_vol.contiguous()
        weights_input = crf.adjust_input_weights(weights, idx)
        affinities_shift = crf.adjust_input_affinities(affinities[:, idx] if affinities is not None else None)
        offsets_shift = crf.adjust_input_offsets(offsets)
        flow, prob_vol, _ = crf.forward(prob_vol, weights_input, affinities_shift, offsets_shift)
    return flow, prob_vol, affinities_shift, offsets_shift


def multiscale_flow_bp(prob_vol_pyramid, crf_pyramid, sws_pyramid, affinity_pyramid=None, unpad=None):
    """Coarse-to-fine flow BP (flow.py:80-115).

    prob_vol_pyramid[l]: [N, 2, H_l, W_l, K_l] (level 0 the finest; H and W halve and
    K_l = (K_{l-1} + 1) / 2 from level to level); crf_pyramid[l]: the BP layers of level l;
    sws_pyramid[l]: the level's search window (K_l = sws_l + 1).  Returns (flow pyramid
    [N, 2, H_l, W_l], beliefs pyramid [N, 2, H_l, W_l, K_l]), finest level first."""
    beliefs_pyramid, flow_pyramid = [], []
    beliefs_in = None
    for lvl in reversed(range(len(prob_vol_pyramid))):
        pv_lvl = prob_vol_pyramid[lvl]
        affinity = affinity_pyramid[lvl] if affinity_pyramid is not None else None
        if beliefs_in is not None:
            N, _, H, W, K = beliefs_in.shape
            size = (2 * H, 2 * W, 2 * K - 1)
            up_u = F.interpolate(beliefs_in[:, 0].unsqueeze(1), size=size, mode='trilinear')[:, 0]
            up_v = F.interpolate(beliefs_in[:, 1].unsqueeze(1), size=size, mode='trilinear')[:, 0]
            beliefs_in = torch.cat((up_u.unsqueeze(1), up_v.unsqueeze(1)), dim=1).contiguous()
            pv_lvl = pv_lvl + beliefs_in / 2.0
        flow_lvl, beliefs_lvl, _, _ = optimize_crf(crf_pyramid[lvl], pv_lvl, None, affinity, None)
        if lvl == 0 and unpad is not None:
            l, r, t, b = unpad
            Hb, Wb = beliefs_lvl.shape[2], beliefs_lvl.shape[3]
            beliefs_lvl = beliefs_lvl[:, :, t:Hb - b, l:Wb - r, :].contiguous()
            flow_lvl = flow_lvl[:, :, t:Hb - b, l:Wb - r].contiguous()
        beliefs_pyramid.append(beliefs_lvl)
        flow_pyramid.append(flow_lvl - sws_pyramid[lvl] // 2)
        beliefs_in = beliefs_pyramid[-1]
    beliefs_pyramid.reverse()
    flow_pyramid.reverse()
    return flow_pyramid, beliefs_pyramid
