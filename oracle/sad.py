"""FP64 restatement of the reference's SAD matching ops (TEST INFRASTRUCTURE ONLY).

Reference sources followed:
* stereo SAD   ops/sad/src/stereo_sad_kernel.cu:16-62 (forward), :64-108 (backward),
               host :116-166; autograd ops/sad/stereo_sad.py:5-19
* flow SAD     ops/flow_mp_sad/src/flow_mp_sad_kernel.cu:10-124 (forward), :126-200
               (backward), host :208-291; autograd ops/flow_mp_sad/flow_mp_sad.py:5-119

Deliberate differences: every batch element is computed (the reference kernels only
touch n=0); the stereo backward covers every channel (the reference grid covers
channels < 4*ceil(D/4) only).

Pinned by tests/test_oracle_sad.py: brute-force loops on tiny inputs written
independently of the vectorised code, and FP64 finite differences of the forward.
"""
import numpy as np

FILL = 40.0          # stereo_sad_kernel.cu:125, flow_mp_sad_kernel.cu:51,217-218
SIGN_EPS = 1e-15     # stereo_sad_kernel.cu:78, flow_mp_sad_kernel.cu:142
FLOW_INIT = 9999999.0  # flow_mp_sad_kernel.cu:86,106


def _sign(t):
    """t/|t| when |t| > 1e-15, else 0 (stereo_sad_kernel.cu:94,101)."""
    return np.where(np.abs(t) > SIGN_EPS, np.sign(t), 0.0)


def stereo_num_disp(min_disp, max_disp, step=1.0):
    """D = (max - min + 1) / step, truncated (stereo_sad_kernel.cu:122)."""
    return int(np.float32(max_disp - min_disp + 1) / np.float32(step))


def stereo_disp(i, min_disp, step=1.0):
    """d = int(i * step + min_disp) (stereo_sad_kernel.cu:35)."""
    return int(np.float32(i) * np.float32(step) + np.float32(min_disp))


def stereo_sad_forward(f0, f1, min_disp, max_disp, step=1.0):
    """cost[n,y,x,i] = sum_c |f0[n,c,y,x] - f1[n,c,y,x-d_i]|, FILL where x-d_i is outside."""
    f0 = np.asarray(f0, dtype=np.float64)
    f1 = np.asarray(f1, dtype=np.float64)
    N, C, H, W = f0.shape
    D = stereo_num_disp(min_disp, max_disp, step)
    out = np.full((N, H, W, D), FILL)
    for i in range(D):
        d = stereo_disp(i, min_disp, step)
        lo, hi = max(d, 0), min(W + d, W)   # x with 0 <= x-d < W
        if lo >= hi:
            continue
        diff = np.abs(f0[:, :, :, lo:hi] - f1[:, :, :, lo - d:hi - d]).sum(axis=1)
        out[:, :, lo:hi, i] = diff
    return out


def stereo_sad_backward(f0, f1, min_disp, max_disp, grad, absolute=False):
    """(df0, df1) of stereo_sad_forward at step 1 (stereo_sad_kernel.cu:64-108).
    ``absolute=True``: the condition magnitudes sum_d |sign| |g| of both outputs."""
    f0 = np.asarray(f0, dtype=np.float64)
    f1 = np.asarray(f1, dtype=np.float64)
    g = np.asarray(grad, dtype=np.float64)
    if absolute:
        g = np.abs(g)
    N, C, H, W = f0.shape
    df0 = np.zeros_like(f0)
    df1 = np.zeros_like(f1)
    for d in range(min_disp, max_disp + 1):
        i = d - min_disp
        lo, hi = max(d, 0), min(W + d, W)
        if lo >= hi:
            continue
        s = _sign(f0[:, :, :, lo:hi] - f1[:, :, :, lo - d:hi - d])  # [N,C,H,w]
        if absolute:
            s = np.abs(s)
        gi = g[:, None, :, lo:hi, i]
        df0[:, :, :, lo:hi] += s * gi
        df1[:, :, :, lo - d:hi - d] += (1.0 if absolute else -1.0) * s * gi
    return df0, df1


def flow_sad_forward(f0, f1, sws, offset_u=0, offset_v=0, block_u=0, block_v=0, want_volume=False):
    """Fused K x K window SAD with min-projections (flow_mp_sad_kernel.cu:10-124).

    K = sws + 1, h = sws // 2.  S(u,v) = sum_c |f0[c,y,x] - f1[c,y+ov+v,x+ou+u]| or FILL.
    u_idx runs over [0, K) but the v loop over [-h, h] (:62-80), so for odd sws the row
    v_idx = K - 1 keeps FILL.
    cv_u[.., u+h] = min_v S, u_star = argmin_v (first) + block_v * sws;
    cv_v[.., v+h] = min_u S, v_star = argmin_u (first) + block_u * sws.
    Argmins are returned as float64 label ids (the reference stores them as fp32).
    ``want_volume``: also return the window volume S [N,H,W,v_idx,u_idx].
    """
    f0 = np.asarray(f0, dtype=np.float64)
    f1 = np.asarray(f1, dtype=np.float64)
    N, C, H, W = f0.shape
    K = sws + 1
    h = sws // 2
    S = np.full((N, H, W, K, K), FILL)  # [.., v_idx, u_idx]
    for vi in range(2 * h + 1):
        dv = offset_v + vi - h
        ylo, yhi = max(0, -dv), min(H, H - dv)
        if ylo >= yhi:
            continue
        for ui in range(K):
            du = offset_u + ui - h
            xlo, xhi = max(0, -du), min(W, W - du)
            if xlo >= xhi:
                continue
            d = np.abs(f0[:, :, ylo:yhi, xlo:xhi] - f1[:, :, ylo + dv:yhi + dv, xlo + du:xhi + du]).sum(axis=1)
            S[:, ylo:yhi, xlo:xhi, vi, ui] = d
    cv_u = S.min(axis=3)
    u_star = S.argmin(axis=3).astype(np.float64) + block_v * sws   # argmin returns first index
    cv_v = S.min(axis=4)
    v_star = S.argmin(axis=4).astype(np.float64) + block_u * sws
    # the reference initialises the running min at 9999999 (only matters if every S >= it)
    if want_volume:
        return cv_u, cv_v, u_star, v_star, S
    return cv_u, cv_v, u_star, v_star


def flow_sad_backward(f0, f1, sws, grad_u, grad_v, u_star, v_star, absolute=False):
    """(df0, df1) routed through the stored argmins (flow_mp_sad_kernel.cu:126-200).

    Labels run over [-h, h] (:152, :175): for odd sws the last label (K - 1) has no gradient.

    For each u: v = u_star[u] - h; if (x+u, y+v) is inside, upd = sign(f0 - f1[y+v,x+u]) *
    grad_u[u]; df0 += upd; df1[y+v, x+u] -= upd.  The same for v with v_star.  Offsets
    and block indices are not applied (as in the reference).  ``absolute=True``: the
    condition magnitudes (|sign| |grad| summed into both outputs)."""
    f0 = np.asarray(f0, dtype=np.float64)
    f1 = np.asarray(f1, dtype=np.float64)
    N, C, H, W = f0.shape
    K = sws + 1
    h = sws // 2
    df0 = np.zeros_like(f0)
    df1 = np.zeros_like(f1)
    ys, xs = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    for n in range(N):
        for k in range(2 * h + 1):
            for gsrc, star, is_u in ((grad_u, u_star, True), (grad_v, v_star, False)):
                other = np.asarray(star[n, :, :, k]).astype(np.int64) - h  # float -> short
                if is_u:
                    u = k - h
                    v = other
                    xx = xs + u
                    yy = ys + v
                else:
                    v = k - h
                    u = other
                    xx = xs + u
                    yy = ys + v
                ok = (xx >= 0) & (xx < W) & (yy >= 0) & (yy < H)
                yy_c = np.clip(yy, 0, H - 1)
                xx_c = np.clip(xx, 0, W - 1)
                diff = f0[n] - f1[n][:, yy_c, xx_c]
                upd = _sign(diff) * np.asarray(gsrc[n, :, :, k], dtype=np.float64)[None] * ok[None]
                if absolute:
                    upd = np.abs(upd)
                df0[n] += upd
                for c in range(C):
                    np.add.at(df1[n, c], (yy_c[ok], xx_c[ok]), (upd if absolute else -upd)[c][ok])
    return df0, df1
