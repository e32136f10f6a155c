"""Parity bars of the GPU tests: fp32 kernels against the FP64 oracle on identical fp32 inputs.

The target is 1e-5 relative error.  "Relative" is measured against the scale that fp32
arithmetic itself is accurate to, which differs by quantity:

* messages (log-probabilities normalised per message vector): per vector,
  max_c |m_gpu - m_64| / max_c |m_64| (``rows``).  The common log-sum-exp shift carries
  the rounding of the largest entries into every entry, so a message entry near 0 has no
  meaningful elementwise relative error;
* beliefs (probabilities): elementwise |b_gpu - b_64| / b_64 (``elementwise``);
* gradients of the LBP op (and of anything downstream of it): normwise over each output
  tensor, max |g_gpu - g_64| / max |g_64| (``normwise``).  The adjoint of a message step
  subtracts the probability-weighted mean (g = t - p sum(t)), so gradient entries are
  differences of much larger sums and carry the rounding of those sums; the absolute-value
  version of that step doubles sum|t| at every pixel of a chain, so componentwise condition
  magnitudes grow like 2^(chain length) and bound nothing.  The elementwise relative error of
  the entries above 10 % of the maximum is printed beside it;
* gradients of single-step sums (the SAD backward, softmax Jacobians): |g_gpu - g_64| / kappa
  elementwise, kappa being the componentwise condition magnitude (the same sums with every
  term by its absolute value, oracle ``absolute=True``); where no term reaches an output
  (kappa = 0) it must be exactly 0 (``conditioned``).

Argmins are compared with the FP64 replay of each step from the kernel's own inputs
(oracle.lbp.local_argmins): identical wherever no competing candidate lies within the fp32
rounding bound of the winner, with the excluded fraction capped.

Every check prints its maximum error, so a run's log holds the numbers.
"""
import numpy as np

TOL = 1e-5


def beliefs_bar(iters):
    """Elementwise bar for beliefs after ``iters`` chained LBP iterations: 2e-5 for one, x3 per
    further iteration.  A belief's relative error is the absolute error of its energy (cost +
    four messages, |E| up to ~20), which carries the rounding accumulated along each message's
    chain; every message of an iteration sums three of the previous one, so it compounds per
    iteration.  Measured: the stereo workload's 4 x 1248 strips at 192 labels err by 3.7e-6 /
    1.2e-5 / 3.5e-5 after 1 / 2 / 3 iterations (the plain fp32 restatement of the reference by
    5.2e-6 / 2.3e-5 / 5.6e-5); the 128 x 256 config-0 image by up to 1.2e-5 after one."""
    return 2.0 * TOL * 3.0 ** (max(1, iters) - 1)
# Decisions within fp32 rounding of a tie are genuine ties at fp32 resolution (for instance
# candidates h + P1 w whose h differ below the ulp of P1 w round to the same value); their
# share is measured and printed per test (up to 0.6 % on the first iteration of a softmax cost
# volume, where most labels have h ~ 0), and must stay below this cap.
ARGMIN_EXCLUDED_CAP = 1e-2


def _show(name, err, tol):
    print(f"    {name}: max err {err:.3g} (bar {tol:.0e})", flush=True)


def rows(name, g, o, tol=TOL, axis=-1):
    """Per-vector relative error along ``axis``; all-zero vectors must match exactly."""
    g = np.asarray(g, dtype=np.float64)
    o = np.asarray(o, dtype=np.float64)
    d = np.abs(g - o).max(axis=axis)
    s = np.abs(o).max(axis=axis)
    assert np.all(d[s == 0] == 0), f"{name}: non-zero values where the oracle's vector is 0"
    err = float((d[s > 0] / s[s > 0]).max()) if np.any(s > 0) else 0.0
    _show(name + " (per vector)", err, tol)
    assert err <= tol, f"{name}: per-vector relative error {err:.3g} > {tol:g}"
    return err


FP32_TINY = 2.0 ** -126   # smallest normal fp32: below it fp32 holds no relative precision


def elementwise(name, g, o, tol=TOL):
    """Elementwise relative error.  Oracle values below the smallest normal fp32 (zero at fp32
    resolution, e.g. an underflowed exp) must come out below it too."""
    g = np.asarray(g, dtype=np.float64)
    o = np.asarray(o, dtype=np.float64)
    z = np.abs(o) < FP32_TINY
    assert np.all(np.abs(g[z]) < FP32_TINY), f"{name}: values where the oracle is 0 at fp32 resolution"
    err = float((np.abs(g - o)[~z] / np.abs(o[~z])).max()) if np.any(~z) else 0.0
    _show(name + " (elementwise)", err, tol)
    assert err <= tol, f"{name}: elementwise relative error {err:.3g} > {tol:g}"
    return err


def normwise(name, g, o, tol=TOL):
    """max |g - o| / max |o| over the whole tensor; an all-zero oracle tensor must be matched
    exactly."""
    g = np.asarray(g, dtype=np.float64)
    o = np.asarray(o, dtype=np.float64)
    s = float(np.abs(o).max()) if o.size else 0.0
    d = np.abs(g - o)
    if s == 0:
        assert np.all(g == 0), f"{name}: non-zero values where the oracle is all 0"
        print(f"    {name}: all zero, matched", flush=True)
        return 0.0
    err = float(d.max()) / s
    big = np.abs(o) >= 0.1 * s
    rel = float((d[big] / np.abs(o[big])).max())
    print(f"    {name}: normwise err {err:.3g} (bar {tol:.0e}); elementwise relative where "
          f"|g| >= 0.1 max: {rel:.3g}", flush=True)
    assert err <= tol, f"{name}: normwise relative error {err:.3g} > {tol:g}"
    return err


def conditioned(name, g, o, kappa, tol=TOL):
    """|g - o| <= tol * kappa elementwise (kappa: condition magnitudes of the oracle)."""
    g = np.asarray(g, dtype=np.float64)
    o = np.asarray(o, dtype=np.float64)
    kappa = np.asarray(kappa, dtype=np.float64)
    z = kappa == 0
    assert np.all(g[z] == 0) and np.all(o[z] == 0), f"{name}: values where no term contributes"
    err = float((np.abs(g - o)[~z] / kappa[~z]).max()) if np.any(~z) else 0.0
    big = ~z & (np.abs(o) >= 0.1 * kappa)   # outputs without strong cancellation
    rel = float((np.abs(g - o)[big] / np.abs(o[big])).max()) if np.any(big) else 0.0
    print(f"    {name}: max |err|/kappa {err:.3g} (bar {tol:.0e}); elementwise relative where "
          f"|g| >= 0.1 kappa: {rel:.3g}", flush=True)
    assert err <= tol, f"{name}: error {err:.3g} x condition magnitude > {tol:g}"
    return err


def argmins(name, gpu_label, gpu_from_trunc, label, from_trunc, clear, written, cap=ARGMIN_EXCLUDED_CAP):
    """GPU argmins (sender label, or the global argmin for truncation winners) equal the FP64
    step replay wherever it is clear of fp32 ties; returns the excluded fraction."""
    wm = np.broadcast_to(np.asarray(written)[..., None], np.shape(clear))
    m = clear & wm
    bad_l = int((np.asarray(gpu_label)[m] != np.asarray(label)[m]).sum())
    bad_t = int((np.asarray(gpu_from_trunc)[m] != np.asarray(from_trunc)[m]).sum())
    excl = float((~clear & wm).sum() / max(1, wm.sum()))
    print(f"    {name}: {bad_l} label / {bad_t} source mismatches; {excl:.2e} of "
          f"{int(wm.sum())} decisions within fp32 rounding of a tie (cap {cap:g})", flush=True)
    assert bad_l == 0 and bad_t == 0, f"{name}: argmins differ from the FP64 step replay"
    assert excl <= cap, f"{name}: excluded fraction {excl:.3g} above {cap:g}"
    return excl


def winner_clear(energy, bound):
    """Pixels whose energy minimum (beliefs' maximum) over the label axis leads the runner-up
    by more than 2 * bound, ``bound`` being the deviation of the GPU's energies from the
    oracle's (a scalar, or per pixel)."""
    e = np.sort(np.asarray(energy, dtype=np.float64), axis=-1)
    return (e[..., 1] - e[..., 0]) > 2.0 * np.asarray(bound)


def energy_deviation(beliefs_gpu, energy):
    """Per pixel, the largest deviation of the GPU's energies, read from its beliefs
    (-log b = E / T + const), from the oracle's energies (both shifted to their minimum)."""
    lg = -np.log(np.maximum(np.asarray(beliefs_gpu, dtype=np.float64), 1e-300))
    E = np.asarray(energy, dtype=np.float64)
    ok = np.asarray(beliefs_gpu) >= FP32_TINY
    d = np.abs((lg - lg.min(-1, keepdims=True)) - (E - E.min(-1, keepdims=True)))
    return np.where(ok, d, 0.0).max(axis=-1)


def loss_grad_of_beliefs(energy, Wb, T=1.0):
    """d/dE of sum(Wb * softmin(E / T)) over the label axis, and its condition magnitudes:
    b = softmin(E / T); dL/dE = -(b / T) (Wb - <Wb, b>); |.| terms: (b / T)(|Wb| + <|Wb|, b>)."""
    z = -np.asarray(energy, dtype=np.float64) / T
    b = np.exp(z - z.max(-1, keepdims=True))
    b /= b.sum(-1, keepdims=True)
    g = -(b / T) * (Wb - (Wb * b).sum(-1, keepdims=True))
    k = (b / T) * (np.abs(Wb) + (np.abs(Wb) * b).sum(-1, keepdims=True))
    return g, k
