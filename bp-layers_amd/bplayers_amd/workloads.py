"""Deterministic synthetic inputs for the benchmark workloads (numpy PCG64, float64 then
cast to float32, so every consumer sees identical bits).

Workloads (sizes as named):
  stereo_fwd_128x256_d64      features [1,32,128,256], disparity field U[0,48], D=64, 1 iteration,
                              global P1=0.1 / P2=0.5, delta 3, forward only
  stereo_fb_384x1248_d192     features [1,64,384,1248], disparity U[0,160], D=192, 3 iterations,
                              per-pixel per-direction P1, P2 = softplus(N(0,1)), forward + backward
  semantic_fb_2x512x1024_c19  logits N(0,2) [2,19,512,1024], 19x19 matrix (zero diagonal, U[0,1]),
                              edges softplus(N(0,1)) [2,4,512,1024], 2 iterations, forward + backward
  flow_fb_256x512_r12         features [1,32,256,512], flow U[-12,12]^2, sws 24 (25 labels per
                              component), P1=0.1 / P2=0.5, 2 iterations, forward + backward
  serving mix                 {256x512 D64, 384x1248 D192, 544x960 D128, 768x1024 D128}, C=64
"""
import numpy as np


def _rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def smooth_field(rng, H, W, lo, hi, grid=(8, 16)):
    """A grid of U[lo,hi] values bilinearly upsampled (align_corners=True) to H x W."""
    gy, gx = grid
    g = rng.uniform(lo, hi, size=(gy, gx))
    ys = np.linspace(0, gy - 1, H)
    xs = np.linspace(0, gx - 1, W)
    y0 = np.clip(np.floor(ys).astype(int), 0, gy - 2)
    x0 = np.clip(np.floor(xs).astype(int), 0, gx - 2)
    wy = (ys - y0)[:, None]
    wx = (xs - x0)[None, :]
    a = g[y0][:, x0]
    b = g[y0][:, x0 + 1]
    c = g[y0 + 1][:, x0]
    d = g[y0 + 1][:, x0 + 1]
    return (1 - wy) * ((1 - wx) * a + wx * b) + wy * ((1 - wx) * c + wx * d)


def warp_x(img, disp):
    """R(y, x') = L(y, x' + d(y, x')): linear interpolation along x, border clamp.  img [C,H,W]."""
    C, H, W = img.shape
    xs = np.arange(W)[None, :] + disp
    xs = np.clip(xs, 0, W - 1)
    x0 = np.clip(np.floor(xs).astype(int), 0, W - 2)
    w = xs - x0
    rows = np.arange(H)[:, None]
    return (1 - w)[None] * img[:, rows, x0] + w[None] * img[:, rows, x0 + 1]


def warp_xy(img, fu, fv):
    """I2(y, x) = I1(y + v, x + u): bilinear, border clamp.  img [C,H,W]."""
    C, H, W = img.shape
    xs = np.clip(np.arange(W)[None, :] + fu, 0, W - 1)
    ys = np.clip(np.arange(H)[:, None] + fv, 0, H - 1)
    x0 = np.clip(np.floor(xs).astype(int), 0, W - 2)
    y0 = np.clip(np.floor(ys).astype(int), 0, H - 2)
    wx = xs - x0
    wy = ys - y0
    a = img[:, y0, x0]
    b = img[:, y0, x0 + 1]
    c = img[:, y0 + 1, x0]
    d = img[:, y0 + 1, x0 + 1]
    return (1 - wy) * ((1 - wx) * a + wx * b) + wy * ((1 - wx) * c + wx * d)


def softplus(x):
    return np.logaddexp(0.0, x)


def stereo_pair(seed, C, H, W, dmax, noise=0.05):
    """Left features N(0,1) [1,C,H,W]; right = left warped by a smooth disparity field in
    [0, dmax] plus N(0, noise).  Returns float32 (f0, f1, disparity field)."""
    rng = _rng(seed)
    left = rng.normal(size=(C, H, W))
    disp = smooth_field(rng, H, W, 0.0, dmax)
    right = warp_x(left, disp) + rng.normal(scale=noise, size=(C, H, W))
    f32 = lambda a: np.ascontiguousarray(a[None], dtype=np.float32)
    return f32(left), f32(right), disp


def stereo_small(seed=0):
    """stereo_fwd_128x256_d64: (f0, f1, D, iters, affinities [1,4,3,H,W])."""
    f0, f1, _ = stereo_pair(seed, 32, 128, 256, 48.0)
    aff = np.zeros((1, 4, 3, 128, 256), dtype=np.float32)
    aff[:, :, :2] = 0.1
    aff[:, :, 2] = 0.5
    return dict(f0=f0, f1=f1, min_disp=0, max_disp=63, iters=1, affinities=aff, delta=3)


def stereo_full(seed=1, H=384, W=1248, C=64, D=192):
    """stereo_fb_384x1248_d192 (or a crop of it with smaller H/W)."""
    f0, f1, _ = stereo_pair(seed, C, H, W, 160.0 * min(1.0, D / 192.0))
    rng = _rng(seed + 1000)
    p1 = softplus(rng.normal(size=(1, 4, H, W)))
    p2 = softplus(rng.normal(size=(1, 4, H, W)))
    aff = np.ascontiguousarray(np.stack([p1, p1, p2], axis=2), dtype=np.float32)
    return dict(f0=f0, f1=f1, min_disp=0, max_disp=D - 1, iters=3, affinities=aff, delta=3)


def semantic(seed=2, N=2, C=19, H=512, W=1024):
    """semantic_fb_2x512x1024_c19: unary prob [N,H,W,C] (softmax of N(0,2) logits),
    pw_mat [1,2,C,C], edge weights [N,4,H,W]."""
    rng = _rng(seed)
    logits = rng.normal(scale=2.0, size=(N, H, W, C))
    logits -= logits.max(axis=-1, keepdims=True)
    prob = np.exp(logits)
    prob /= prob.sum(axis=-1, keepdims=True)
    M = rng.uniform(0.0, 1.0, size=(C, C))
    np.fill_diagonal(M, 0.0)
    edge = softplus(rng.normal(size=(N, 4, H, W)))
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)
    return dict(prob=f32(prob), pw_mat=f32(np.stack([M, M])[None]), edge=f32(edge), iters=2, delta=C)


def flow(seed=3, C=32, H=256, W=512, R=12):
    """flow_fb_256x512_r12: frames [1,C,H,W] with frame 2 = frame 1 warped by a smooth flow."""
    rng = _rng(seed)
    f1 = rng.normal(size=(C, H, W))
    fu = smooth_field(rng, H, W, -R, R)
    fv = smooth_field(rng, H, W, -R, R)
    f2 = warp_xy(f1, fu, fv) + rng.normal(scale=0.05, size=(C, H, W))
    f32 = lambda a: np.ascontiguousarray(a[None], dtype=np.float32)
    return dict(f0=f32(f1), f1=f32(f2), sws=2 * R, iters=2, P1=0.1, P2=0.5, delta=3)


SERVING_SHAPES = ((256, 512, 64), (384, 1248, 192), (544, 960, 128), (768, 1024, 128))


def serving_request(seed, shape_idx, C=64):
    """One serving request: a stereo pair of the given shape class with C=64 features."""
    H, W, D = SERVING_SHAPES[shape_idx]
    f0, f1, _ = stereo_pair(seed, C, H, W, 0.75 * D)
    return dict(f0=f0, f1=f1, max_disp=D - 1, H=H, W=W, D=D)
