"""Multi-scale stereo BP around the ops (reference stereo.py:95-133 and 180-197): per pyramid
level, the BP layers run on the level's probability volume, coarse to fine; the beliefs of a
level are upsampled x2 in height, width and disparity (trilinear) and half of them added to
the next finer level's probabilities before its BP.

Behaviour kept from the reference:
  * the beliefs are added only when they carry no gradient (inference), and in place (the
    caller's probability pyramid changes); with gradients the finer level's probabilities are
    renormalised over the disparities instead and the coarser beliefs are not used;
  * each level's winner-take-all disparities are offset by the level's ``min_disp``;
  * the returned pyramids are reordered after every level (the reference reverses its lists
    inside the level loop), so entry 0 is the finest level; for three levels the order is
    [level 0, level 2, level 1];
  * with ``unpad`` = (l, r, t, b) the finest level's beliefs (NHWC) and disparities (NCHW) are
    cropped as the reference crops its padded input."""
import torch
import torch.nn.functional as F

from .corenet import crop_padding


def optimize_crf(crf_layers, prob_vol, weights, affinities):
    """Runs the BP layers of one level in turn (stereo.py:180-197): each layer gets the
    adjusted edge weights, affinities and offsets and the previous layer's beliefs as its
    probability volume.  Returns (disparities, beliefs, adjusted affinities, adjusted offsets),
    or None without layers."""
    if not crf_layers:
        return None
    offsets = None
    disps = affinities_shift = offsets_shift = None
    for idx, crf in enumerate(crf_layers):
        prob_vol = prob_vol.contiguous()
        weights_input = crf.adjust_input_weights(weights, idx)
        affinities_shift = crf.adjust_input_affinities(affinities)
        offsets_shift = crf.adjust_input_offsets(offsets)
        disps, prob_vol, _ = crf.forward(prob_vol, weights_input, affinities_shift, offsets_shift)
    return disps, prob_vol, affinities_shift, offsets_shift


def multiscale_bp(prob_vol_pyramid, crf_pyramid, affinity_pyramid=None, min_disps=None, unpad=None):
    """Coarse-to-fine BP over a pyramid (stereo.py:95-133).

    prob_vol_pyramid[l]: [N, H_l, W_l, D_l] probabilities, level 0 the finest, each level half
    the previous one in H, W and D; crf_pyramid[l]: the list of BP layers of level l;
    affinity_pyramid[l]: [1, 2, 5, H_l, W_l] or None; min_disps[l]: the level's minimum
    disparity (0 by default).  Returns (disparity pyramid, beliefs pyramid) in the reference's
    order (entry 0 the finest level)."""
    levels = len(prob_vol_pyramid)
    min_disps = min_disps if min_disps is not None else [0] * levels
    beliefs_pyramid, disps_pyramid = [], []
    beliefs_in = None
    for lvl in reversed(range(levels)):
        pv_lvl = prob_vol_pyramid[lvl]
        affinity = affinity_pyramid[lvl] if affinity_pyramid is not None else None
        if beliefs_in is not None:
            beliefs_in = F.interpolate(beliefs_in.unsqueeze(1), scale_factor=2.0, mode='trilinear')[:, 0]
            if beliefs_in.requires_grad:
                pv_lvl = pv_lvl / pv_lvl.sum(dim=-1, keepdim=True)
            else:
                pv_lvl += beliefs_in / 2.0
            beliefs_in = None
        disps_lvl, beliefs_lvl, _, _ = optimize_crf(crf_pyramid[lvl], pv_lvl, None, affinity)
        if lvl == 0 and unpad is not None:
            l, r, t, b = unpad
            beliefs_lvl = crop_padding(beliefs_lvl.permute(0, 3, 1, 2), (l, r, t, b)).permute(0, 2, 3, 1).contiguous()
            disps_lvl = crop_padding(disps_lvl, (l, r, t, b))
        beliefs_pyramid.append(beliefs_lvl)
        disps_pyramid.append(disps_lvl + min_disps[lvl])
        beliefs_in = beliefs_pyramid[-1]
        beliefs_pyramid.reverse()
        disps_pyramid.reverse()
    return disps_pyramid, beliefs_pyramid
