#!/usr/bin/env python3
"""Benchmark of the BP-Layers hot path on one MI355X.

Default workload: stereo_fb_384x1248_d192 -- full-resolution stereo LBP, batch 1, 192
disparities, 3 iterations with per-pixel, per-direction P1/P2, forward + exact backward.
A step is one pass of the LBP hot path over the batch: 3 forward iterations followed by
their 3 backward iterations, on inputs already resident in HBM (the SAD cost volume is
computed once before timing and timed separately).

Prints ONE JSON line (see the contract in the repository docs):
  metric "G(pixel*label)/s per LBP iter fwd+bwd", value = N*H*W*C*iters / t_step / 1e9.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "bp-layers_amd"))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published
# canonical algorithmic bytes per pixel*label per LBP iteration (fp32 messages, u8 argmins,
# passes fused per phase, chain message on chip): forward 44, exact backward 52; per-pixel
# parameters add (P1/P2 per direction: 32 B/px read fwd, 64 B/px read+grad bwd; edge weights
# 16 / 32 B/px) / C.  Split by phase (kernel class of libbpl_amd):
#   forward horizontal  read cost, M_U, M_D; write M_L, M_R, 2 argmins        22 B/pl
#   forward vertical    read cost, M_L, M_R; write M_U, M_D, 2 argmins        22 B/pl
#   backward vertical   read M_U, M_D, 2 argmins, G_U, G_D; write S_V         22 B/pl
#   backward horizontal read M_L, M_R, 2 argmins, G_L, G_R, S_V;
#                       write g_cost, g_msg (with the combine kernel)         30 B/pl
FWD_B_PL = 44.0
BWD_B_PL = 52.0
ACC_B_PL = 4.0  # summing grad_cost over a chain's iterations: read the running sum
PHASE_B_PL = {"lbp_fwd_horizontal": (22.0, 24.0), "lbp_fwd_vertical": (22.0, 24.0),
              "lbp_bwd_vertical": (22.0, 48.0), "lbp_bwd_horizontal": (30.0, 48.0)}  # (per pl, per px / C)
LBP_CLASSES = ("lbp_fwd_horizontal", "lbp_fwd_vertical", "lbp_bwd_vertical", "lbp_bwd_horizontal",
               "lbp_bwd_combine")
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic.json")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="stereo_fb_384x1248_d192",
                    choices=["stereo_fb_384x1248_d192", "stereo_fwd_128x256_d64", "semantic_fb_2x512x1024_c19",
                             "semantic_pixel_fb_2x512x1024_c19", "flow_fb_256x512_r12", "serving_c4",
                             "serving_batch"])
    ap.add_argument("--requests", type=int, default=96, help="serving_c4: requests per measurement")
    ap.add_argument("--rows", type=int, default=0, help="crop the image to this many rows (debug)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-others", action="store_true",
                    help="default workload: skip the compact semantic / flow / serving results")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--pmc-summary", nargs="+", metavar="DIR",
                    help="summarise rocprofv3 --pmc runs of `bench.py --pmc` (directories holding "
                         "*counter_collection.csv) into profiles/traffic.json and print a table")
    ap.add_argument("--pmc", action="store_true",
                    help="counter-collection run for rocprofv3 --pmc: one warm-up step, one step and a "
                         "256 MiB device copy (the calibration kernel); prints what ran")
    return ap.parse_args()


class StereoPlan:
    """All device buffers of the LBP fwd+bwd step, allocated once."""

    def __init__(self, wl, iters, rows=0, backward=True):
        import numpy as np
        import torch
        from bplayers_amd import _native as nat
        from bplayers_amd import pytorch_cuda_stereo_sad_op as sad
        from bplayers_amd.ops.lbp_stereo.message_passing_op_cuda import pack_affinities
        self.nat = nat
        dev = torch.device("cuda:0")
        self.dev = dev
        f0 = torch.from_numpy(wl["f0"]).to(dev)
        f1 = torch.from_numpy(wl["f1"]).to(dev)
        if rows:
            f0 = f0[:, :, :rows].contiguous()
            f1 = f1[:, :, :rows].contiguous()
        self.f0, self.f1 = f0, f1
        self.min_disp, self.max_disp = wl["min_disp"], wl["max_disp"]
        sadv = sad.forward(f0, f1, self.min_disp, self.max_disp, 1.0)
        self.cost = (-torch.softmax(-sadv, dim=3)).contiguous()     # cost = -softmax(-SAD), T = 1
        self.sad_grad = torch.randn_like(sadv)
        del sadv
        N, H, W, C = self.cost.shape
        self.N, self.H, self.W, self.C = N, H, W, C
        aff = torch.from_numpy(wl["affinities"]).to(dev)
        if rows:
            aff = aff[..., :rows, :].contiguous()
        self.jump = pack_affinities(aff, torch.zeros((N, 4, H, W), device=dev))
        self.K = self.jump.shape[2]
        self.edge = torch.ones((N, 4, H, W), device=dev)
        self.delta = wl["delta"]
        self.iters = iters
        self.msgs = [torch.zeros((N, 4, H, W, C), device=dev)] + \
                    [torch.empty((N, 4, H, W, C), device=dev) for _ in range(iters)]
        self.argmin = [torch.empty((N, 4, H, W, C), dtype=torch.uint8, device=dev) for _ in range(iters)]
        self.garg = [torch.empty((N, 4, H, W), dtype=torch.uint8, device=dev) for _ in range(iters)]
        self.scale = [torch.empty((N, 4, H, W), device=dev) for _ in range(iters)]  # message_scale
        self.backward = backward
        if backward:
            g = torch.Generator(device="cpu").manual_seed(5)
            self.G = torch.randn((N, 4, H, W, C), generator=g).to(dev)
            self.gmsg = [torch.empty((N, 4, H, W, C), device=dev) for _ in range(2)]
            self.gcost = torch.empty((N, H, W, C), device=dev)
            self.gjump = torch.empty_like(self.jump)
            self.gedge = torch.empty((N, 4, H, W), device=dev)
            ws = int(nat.lib().bpl_lbp_bwd_workspace_size(nat.LBP_STEREO, N, H, W, C))
            self.ws = torch.empty(ws, dtype=torch.uint8, device=dev)
            self.ws_bytes = ws
        self.fws_bytes = int(nat.lib().bpl_lbp_fwd_workspace_size(nat.LBP_STEREO, N, H, W, C))
        self.fws = torch.empty(self.fws_bytes, dtype=torch.uint8, device=dev)
        self.stream = torch.cuda.current_stream(dev)
        self.sp = self.nat.ptr  # shorthand

    def pl(self):
        return self.N * self.H * self.W * self.C

    def fwd(self, t):
        nat, p = self.nat, self.sp
        # the workspace form: the vertical phase reads the horizontal phase's base plane
        nat.call("bpl_lbp_fwd_ws", nat.LBP_STEREO, p(self.cost), p(self.jump), self.K, self.jump.shape[0],
                 p(self.edge), p(self.msgs[t]), p(self.msgs[t + 1]), p(self.argmin[t]), p(self.garg[t]),
                 p(self.scale[t]), self.N, self.H, self.W, self.C, self.delta, p(self.fws), self.fws_bytes,
                 nat.FWD_MSG_IN_ZERO if t == 0 else 0, nat.stream_of(self.cost))

    def bwd(self, t, g_in, g_out):
        nat, p = self.nat, self.sp
        # chained iterations: the message gradients between two backward calls carry their
        # LEFT/RIGHT slots at the row ends only (zero elsewhere in the exact adjoint); the top
        # call reads the full loss gradient; the bottom one computes no gradient w.r.t. the
        # initial messages, constant zeros that nothing trains (as BP's); the calls after the
        # top one add their cost, pairwise and edge gradients to the top one's (the sum over
        # iterations of a training step)
        top = t == self.iters - 1
        mode = nat.GRAD_EXACT | (0 if top else nat.GRAD_IN_H_ROW_ENDS | nat.GRAD_IN_UD_SHARED | nat.GRAD_ACCUMULATE) | \
            (nat.GRAD_OUT_H_ROW_ENDS | nat.GRAD_OUT_UD_SHARED if t > 0 else 0)
        nat.call("bpl_lbp_bwd", nat.LBP_STEREO, mode, p(self.jump), self.K, self.jump.shape[0],
                 p(self.edge), p(g_in), p(self.msgs[t + 1]), p(self.argmin[t]), p(self.garg[t]), p(self.gcost),
                 p(self.gjump), p(self.gedge), p(g_out) if t > 0 else None, p(self.ws), self.ws_bytes, self.N,
                 self.H, self.W, self.C, self.delta, nat.stream_of(self.cost))

    def step(self):
        for t in range(self.iters):
            self.fwd(t)
        if self.backward:
            g_in = self.G
            for t in reversed(range(self.iters)):
                g_out = self.gmsg[t % 2]
                self.bwd(t, g_in, g_out)
                g_in = g_out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(wl, plan_shape, seconds, iters_per_sample=1):
    """The C++/OpenMP restatement (oracle/c) on a crop of the same workload, fp32, all host
    cores: forward + exact backward of one LBP iteration; returns G pl/s, with a single-thread
    figure on a smaller crop (about a third of the time budget) and the CPU model."""
    import numpy as np
    sys.path.insert(0, ROOT)
    from oracle import cpu
    N, H, W, C = plan_shape
    all_threads = len(os.sched_getaffinity(0))
    rng = np.random.default_rng(0)

    def run(rows, threads):
        cost = -rng.uniform(0, 1, size=(N, rows, W, C)).astype(np.float32)
        aff = wl["affinities"][..., :rows, :]
        from bplayers_amd.ops.lbp_stereo.message_passing_op_cuda import pack_affinities
        import torch
        jump = pack_affinities(torch.from_numpy(np.ascontiguousarray(aff)), torch.zeros((N, 4, rows, W))).numpy()
        edge = np.ones((N, 4, rows, W), dtype=np.float32)
        M0 = np.zeros((N, 4, rows, W, C), dtype=np.float32)
        G = rng.normal(size=M0.shape).astype(np.float32)
        t0 = time.perf_counter()
        M, am, fg = cpu.lbp_forward(cost, jump, edge, M0, wl["delta"], threads=threads)
        cpu.lbp_backward_exact(jump, edge, G, M, am, fg, wl["delta"], threads=threads)
        return time.perf_counter() - t0

    def sized(threads, budget):
        rows = 4
        t = run(rows, threads)
        # scale the crop so the timed sample takes about `budget`
        rows = int(max(4, min(H, rows * budget / max(t, 1e-3))))
        t = run(rows, threads)
        # the whole image takes less than the budget: repeat it (fresh inputs each time) and average
        reps = max(1, int(budget / max(t, 1e-3))) if rows == H else 1
        if reps > 1:
            t = sum(run(rows, threads) for _ in range(reps)) / reps
        return rows, t, reps

    rows, t, reps = sized(all_threads, seconds)
    pl = N * rows * W * C
    rows1, t1, _ = sized(1, seconds / 3)
    pl1 = N * rows1 * W * C
    return {"value": pl / t / 1e9, "unit": "G(pixel*label)/s per LBP iter fwd+bwd", "cores": all_threads,
            "kind": "port", "cpu_model": cpu_model(),
            "sample": f"C++/OpenMP fp32 restatement (oracle/c), 1 iteration fwd+exact bwd on a "
                      f"{rows}x{W} image with {C} labels ({pl / 1e6:.1f} M pl), mean of {reps} runs "
                      f"of {t:.2f} s ({reps * t:.0f} s of CPU time)",
            "single_thread": {"value": pl1 / t1 / 1e9, "cores": 1,
                              "sample": f"the same on a {rows1}x{W} crop, one thread, {t1:.2f} s"}}


class LBPPlan:
    """Buffers of an LBP step for any pairwise variant: ``iters`` forward iterations from the
    given messages, then (``backward``) their exact backward from a fixed synthetic gradient."""

    def __init__(self, kind, cost, jump, K, Nj, edge, delta, iters, backward=True, msgs0=None):
        import torch
        from bplayers_amd import _native as nat
        self.nat, self.sp = nat, nat.ptr
        self.kind, self.cost, self.jump, self.K, self.Nj, self.edge = kind, cost, jump, K, Nj, edge
        self.delta, self.iters, self.backward = delta, iters, backward
        N, H, W, C = cost.shape
        self.N, self.H, self.W, self.C = N, H, W, C
        dev = cost.device
        m0 = msgs0 if msgs0 is not None else torch.zeros((N, 4, H, W, C), device=dev)
        self.msgs = [m0] + [torch.empty((N, 4, H, W, C), device=dev) for _ in range(iters)]
        self.m0_zero = msgs0 is None
        self.argmin = [torch.empty((N, 4, H, W, C), dtype=torch.uint8, device=dev) for _ in range(iters)]
        self.garg = [torch.empty((N, 4, H, W), dtype=torch.uint8, device=dev) for _ in range(iters)]
        self.scale = [torch.empty((N, 4, H, W), device=dev) for _ in range(iters)]
        self.fws_bytes = int(nat.lib().bpl_lbp_fwd_workspace_size(kind, N, H, W, C))
        self.fws = torch.empty(self.fws_bytes, dtype=torch.uint8, device=dev) if self.fws_bytes else None
        if backward:
            g = torch.Generator(device="cpu").manual_seed(5)
            self.G = torch.randn((N, 4, H, W, C), generator=g).to(dev)
            self.gmsg = [torch.empty((N, 4, H, W, C), device=dev) for _ in range(2)]
            self.gcost = torch.empty((N, H, W, C), device=dev)
            self.gjump = torch.empty_like(jump)
            self.gedge = torch.empty((N, 4, H, W), device=dev)
            self.ws_bytes = int(nat.lib().bpl_lbp_bwd_workspace_size(kind, N, H, W, C))
            self.ws = torch.empty(max(self.ws_bytes, 1), dtype=torch.uint8, device=dev)

    def pl(self):
        return self.N * self.H * self.W * self.C

    def fwd(self, t):
        nat, p = self.nat, self.sp
        nat.call("bpl_lbp_fwd_ws", self.kind, p(self.cost), p(self.jump), self.K, self.Nj, p(self.edge),
                 p(self.msgs[t]), p(self.msgs[t + 1]), p(self.argmin[t]), p(self.garg[t]), p(self.scale[t]),
                 self.N, self.H, self.W, self.C, self.delta, p(self.fws), self.fws_bytes,
                 nat.FWD_MSG_IN_ZERO if t == 0 and self.m0_zero else 0, nat.stream_of(self.cost))

    def bwd(self, t, g_in, g_out):
        nat, p = self.nat, self.sp
        # chained iterations (see StereoPlan.bwd): from zero initial messages the bottom call
        # computes no gradient w.r.t. them; the per-pixel variant's pairwise gradient is not
        # accumulated by the library
        top = t == self.iters - 1
        acc = 0 if top or self.kind == nat.LBP_SEMANTIC_PIXEL else nat.GRAD_ACCUMULATE
        mode = nat.GRAD_EXACT | (0 if top else nat.GRAD_IN_H_ROW_ENDS | nat.GRAD_IN_UD_SHARED) | acc | \
            (nat.GRAD_OUT_H_ROW_ENDS | nat.GRAD_OUT_UD_SHARED if t > 0 else 0)
        nat.call("bpl_lbp_bwd", self.kind, mode, p(self.jump), self.K, self.Nj, p(self.edge), p(g_in),
                 p(self.msgs[t + 1]), p(self.argmin[t]), p(self.garg[t]), p(self.gcost), p(self.gjump),
                 p(self.gedge), p(g_out) if t > 0 or not self.m0_zero else None, p(self.ws), self.ws_bytes,
                 self.N, self.H, self.W, self.C, self.delta, nat.stream_of(self.cost))

    def step(self):
        for t in range(self.iters):
            self.fwd(t)
        if self.backward:
            g_in = self.G
            for t in reversed(range(self.iters)):
                g_out = self.gmsg[t % 2]
                self.bwd(t, g_in, g_out)
                g_in = g_out


def timed_steps(steps, warmup, fn, nat, torch):
    """Run ``fn`` warmup + steps times; returns (ms per step, wall s, {kernel: (ms, launches)})."""
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    nat.timing_reset()
    nat.timing_enable(True)
    t0 = time.perf_counter()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(steps):
        fn()
    end.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    nat.timing_enable(False)
    return start.elapsed_time(end) / steps, wall, nat.timing_read()


def common_line(args, metric, value, unit, ms_step, config, dtype="f32", higher=True):
    return {"metric": metric, "value": round(value, 4), "unit": unit, "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(ms_step, 4), "higher_is_better": higher, "scaling": "weak",
            "vs_baseline": None, "dtype": dtype, "data": "synthetic", "config": config}


def bench_semantic(args, torch):
    """semantic_fb_2x512x1024_c19: N=2, 512x1024, 19 labels, global C x C pairwise (pw_mat used for
    both directions' pairs, RIGHT/DOWN transposed), per-pixel edge weights, delta = C (full
    min-plus), 2 iterations forward + exact backward per step."""
    from bplayers_amd import workloads, _native as nat
    from bplayers_amd.ops.lbp_semantic_pw.message_passing_op_cuda import pack_pairwise
    wl = workloads.semantic()
    dev = torch.device("cuda:0")
    cost = (-torch.from_numpy(wl["prob"])).to(dev).contiguous()
    jump = pack_pairwise(torch.from_numpy(wl["pw_mat"]).to(dev), False)
    edge = torch.from_numpy(wl["edge"]).to(dev)
    plan = LBPPlan(nat.LBP_SEMANTIC_GLOBAL, cost, jump, 0, 1, edge, wl["delta"], wl["iters"])
    if args.pmc:
        return pmc_steps(plan.step, torch, dev, {"workload": args.workload, "pl": plan.pl()})
    ms, wall, kt = timed_steps(args.steps, args.warmup, plan.step, nat, torch)
    value = plan.pl() * plan.iters / (ms * 1e-3) / 1e9
    avg = {k: v[0] / v[1] for k, v in kt.items()}
    it_ms = sum(avg[k] for k in LBP_CLASSES if k in avg) + avg.get("lbp_bwd_reduce", 0.0)
    C = plan.C
    it_bytes = plan.pl() * (FWD_B_PL + BWD_B_PL + (16.0 + 32.0) / C)  # edge weights per pixel
    line = common_line(args, "G(pixel*label)/s per LBP iter fwd+bwd", value, "G(pixel*label)/s", ms,
                       {"workload": args.workload, "N": plan.N, "H": plan.H, "W": plan.W, "labels": C,
                        "iters": plan.iters, "delta": plan.delta, "pairwise": "global CxC matrix per direction",
                        "grad_mode": "exact"})
    line["roofline_iteration"] = {"bound": "hbm", "kernels": [k for k in avg if k.startswith("lbp")],
                                  "achieved": round(it_bytes / (it_ms * 1e-3) / 1e9, 2), "peak": HBM_PEAK_GBS,
                                  "unit": "GB/s", "frac": round(it_bytes / (it_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                                  "alg_bytes_per_pl": round(it_bytes / plan.pl(), 3), "iteration_ms": round(it_ms, 4)}
    line["timings_ms"] = {f"{k}_ms": round(v, 4) for k, v in avg.items()}
    line["wall_s"] = round(wall, 3)
    return line


def bench_semantic_pixel(args, torch):
    """semantic_pixel_fb_2x512x1024_c19: the pixel-wise variant of the semantic workload -- per
    pixel and direction the 19x19 matrix scaled by softplus(N(0,1)) weights (of image 0; the
    op's matrices [4,H,W,C,C] are shared by the batch, 3 GB), unit edge weights, delta = C,
    2 iterations forward + exact backward per step."""
    from bplayers_amd import workloads, _native as nat
    wl = workloads.semantic()
    dev = torch.device("cuda:0")
    cost = (-torch.from_numpy(wl["prob"])).to(dev).contiguous()
    N, H, W, C = cost.shape
    M = torch.from_numpy(wl["pw_mat"][0, 0]).to(dev)
    w = torch.from_numpy(wl["edge"][0]).to(dev)  # [4,H,W]
    jump = (w[:, :, :, None, None] * M).contiguous()
    edge = torch.ones((N, 4, H, W), device=dev)
    plan = LBPPlan(nat.LBP_SEMANTIC_PIXEL, cost, jump, 0, 1, edge, wl["delta"], wl["iters"])
    if args.pmc:
        return pmc_steps(plan.step, torch, dev, {"workload": args.workload, "pl": plan.pl()})
    ms, wall, kt = timed_steps(args.steps, args.warmup, plan.step, nat, torch)
    value = plan.pl() * plan.iters / (ms * 1e-3) / 1e9
    avg = {k: v[0] / v[1] for k, v in kt.items()}
    it_ms = sum(avg[k] for k in LBP_CLASSES if k in avg) + avg.get("lbp_bwd_reduce", 0.0)
    # per pixel and direction the sender's C x C matrix is read in the forward and read and
    # written (its gradient) in the backward, once per image of the batch
    it_bytes = plan.pl() * (FWD_B_PL + BWD_B_PL) + 4 * H * W * C * C * 4 * (N + 2 * N)
    line = common_line(args, "G(pixel*label)/s per LBP iter fwd+bwd", value, "G(pixel*label)/s", ms,
                       {"workload": args.workload, "N": N, "H": H, "W": W, "labels": C, "iters": plan.iters,
                        "delta": plan.delta, "pairwise": "per-pixel CxC matrix per direction", "grad_mode": "exact"})
    line["roofline_iteration"] = {"bound": "hbm", "kernels": [k for k in avg if k.startswith("lbp")],
                                  "achieved": round(it_bytes / (it_ms * 1e-3) / 1e9, 2), "peak": HBM_PEAK_GBS,
                                  "unit": "GB/s", "frac": round(it_bytes / (it_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                                  "alg_bytes_per_pl": round(it_bytes / plan.pl(), 3), "iteration_ms": round(it_ms, 4)}
    line["timings_ms"] = {f"{k}_ms": round(v, 4) for k, v in avg.items()}
    line["wall_s"] = round(wall, 3)
    return line


def bench_flow(args, torch):
    """flow_fb_256x512_r12: flow SAD over a 25x25 window (sws=24) with min-projections to u and
    v volumes [1,256,512,25], cost = -softmax(-SAD), then 2 LBP iterations on u and 2 on v
    (v starting from u's messages, as the reference's 5-D BP path), forward + exact backward,
    per-pixel-constant P1=0.1 / P2=0.5."""
    from bplayers_amd import workloads, _native as nat
    from bplayers_amd.ops.lbp_stereo.message_passing_op_cuda import pack_affinities
    wl = workloads.flow()
    dev = torch.device("cuda:0")
    f0 = torch.from_numpy(wl["f0"]).to(dev)
    f1 = torch.from_numpy(wl["f1"]).to(dev)
    N, Cf, H, W = f0.shape
    K = wl["sws"] + 1
    cu, cv, us, vs = (torch.empty((N, H, W, K), device=dev) for _ in range(4))
    st = nat.stream_of(f0)
    sad_args = (nat.ptr(f0), nat.ptr(f1), nat.ptr(cu), nat.ptr(cv), nat.ptr(us), nat.ptr(vs), N, Cf, H, W, wl["sws"],
                0, 0, 0, 0, st)
    nat.call("bpl_sad_flow_fwd", *sad_args)
    cost_u = (-torch.softmax(-cu, dim=3)).contiguous()
    cost_v = (-torch.softmax(-cv, dim=3)).contiguous()
    aff = torch.zeros((1, 4, 3, H, W), device=dev)
    aff[:, :, :2] = wl["P1"]
    aff[:, :, 2] = wl["P2"]
    jump = pack_affinities(aff, torch.zeros((1, 4, H, W), device=dev))
    edge = torch.ones((N, 4, H, W), device=dev)
    pu = LBPPlan(nat.LBP_STEREO, cost_u, jump, jump.shape[2], 1, edge, wl["delta"], wl["iters"])
    pv = LBPPlan(nat.LBP_STEREO, cost_v, jump, jump.shape[2], 1, edge, wl["delta"], wl["iters"],
                 msgs0=pu.msgs[-1])

    def step():
        for t in range(pu.iters):
            pu.fwd(t)
        for t in range(pv.iters):
            pv.fwd(t)
        for plan in (pv, pu):
            g_in = plan.G
            for t in reversed(range(plan.iters)):
                g_out = plan.gmsg[t % 2]
                plan.bwd(t, g_in, g_out)
                g_in = g_out

    df0, df1 = torch.empty_like(f0), torch.empty_like(f1)
    gu, gv = torch.randn_like(cu), torch.randn_like(cv)

    def sad_step():
        nat.call("bpl_sad_flow_fwd", *sad_args)
        nat.call("bpl_sad_flow_bwd", nat.ptr(f0), nat.ptr(f1), nat.ptr(gu), nat.ptr(gv), nat.ptr(us), nat.ptr(vs),
                 nat.ptr(df0), nat.ptr(df1), N, Cf, H, W, wl["sws"], st)

    if args.pmc:
        def both():
            sad_step()
            step()
        return pmc_steps(both, torch, dev, {"workload": args.workload, "pl": pu.pl() + pv.pl()})
    ms, wall, kt = timed_steps(args.steps, args.warmup, step, nat, torch)
    pl_step = (pu.pl() + pv.pl()) * wl["iters"]
    avg = {k: v[0] / v[1] for k, v in kt.items()}
    # flow SAD forward and backward, timed the same way
    _, _, kt2 = timed_steps(5, 1, sad_step, nat, torch)
    for k, v in kt2.items():
        avg[k] = v[0] / v[1]
    line = common_line(args, "G(pixel*label)/s per LBP iter fwd+bwd", pl_step / (ms * 1e-3) / 1e9,
                       "G(pixel*label)/s", ms,
                       {"workload": args.workload, "N": N, "H": H, "W": W, "labels": K, "iters": wl["iters"],
                        "volumes": "u and v", "delta": wl["delta"], "pairwise": "global P1=0.1 P2=0.5 (K=6 jump)",
                        "grad_mode": "exact"})
    line["timings_ms"] = {f"{k}_ms": round(v, 4) for k, v in avg.items()}
    line["sad"] = {"flow_fwd_G_pixel_window_per_s": round(N * H * W * K * K / (avg["sad_flow_fwd"] * 1e-3) / 1e9, 3),
                   "window": f"{K}x{K}", "feature_channels": Cf}
    line["wall_s"] = round(wall, 3)
    return line


def bench_serving(args, torch):
    """serving_c4: stereo requests (C=64 features) of four shape classes {256x512 D64,
    384x1248 D192, 544x960 D128, 768x1024 D128}, each solved forward-only (SAD -> softmax ->
    1 LBP iteration -> beliefs / WTA disparity). Capacity = frames/s with every request queued
    at t=0 (continuous batching vs one request at a time, same kernels); latency p50/p99 at a
    Poisson offered load of 70% of the sequential capacity."""
    from bplayers_amd import workloads
    from bplayers_amd.serving import StereoServer, make_requests
    dev = torch.device("cuda:0")
    pool = {}
    for i, shp in enumerate(workloads.SERVING_SHAPES):
        r = workloads.serving_request(40 + i, i)
        pool[shp] = [(torch.from_numpy(r["f0"]).to(dev), torch.from_numpy(r["f1"]).to(dev))]
    n = args.requests
    budget = 96 << 30

    def run(sequential, rate, seed):
        srv = StereoServer(dev, C=64, hbm_budget=budget, sequential=sequential, max_batch=4, max_slots_per_class=2)
        srv.warm(sorted(pool))
        srv.serve(make_requests(pool, 8, rate_hz=1e9, seed=99))  # warm-up (not measured)
        return srv.serve(make_requests(pool, n, rate_hz=rate, seed=seed))

    seq = run(True, 1e9, 1)
    mix = run(False, 1e9, 1)
    rate = 0.7 * seq["frames_per_s"]
    seq_l = run(True, rate, 2)
    mix_l = run(False, rate, 2)
    line = common_line(args, "frames/s (stereo serving, 4 shape classes)", mix["frames_per_s"], "frames/s",
                       1e3 / mix["frames_per_s"],
                       {"workload": args.workload, "shapes": [list(s) for s in workloads.SERVING_SHAPES],
                        "feature_channels": 64, "iters": 1, "requests": n, "max_batch": 4, "slots_per_class": 2})
    line["sequential_frames_per_s"] = round(seq["frames_per_s"], 3)
    line["speedup_vs_sequential"] = round(mix["frames_per_s"] / seq["frames_per_s"], 3)
    line["latency_ms_at_offered_load"] = {"offered_frames_per_s": round(rate, 3),
                                          "continuous": {"p50": round(mix_l["p50_ms"], 2), "p99": round(mix_l["p99_ms"], 2)},
                                          "sequential": {"p50": round(seq_l["p50_ms"], 2), "p99": round(seq_l["p99_ms"], 2)}}
    line["resident_bytes"] = mix["resident_bytes"]
    return line


def bench_serving_batch(args, torch):
    """serving_batch: GPU time per frame of one serving slot (SAD + softmax, 1 LBP iteration,
    beliefs; C = 64 features, forward only, messages without argmins) at batch 1, 2 and 4 for
    each shape class: ``args.steps`` back-to-back launches timed with events, and the library's
    per-kernel times.  With --pmc: one launch of the 384x1248 D192 class at batch 1 and one at
    batch 4 (for rocprofv3 counter passes; dispatches of the two differ in grid size)."""
    from bplayers_amd import workloads
    from bplayers_amd import _native as nat
    from bplayers_amd.serving import Slot, Request
    dev = torch.device("cuda:0")
    shapes = list(workloads.SERVING_SHAPES)
    if args.pmc:
        shapes = [(384, 1248, 192)]
    out = {}
    for i, shp in enumerate(shapes):
        r = workloads.serving_request(40 + i, workloads.SERVING_SHAPES.index(shp))
        f0, f1 = torch.from_numpy(r["f0"]).to(dev), torch.from_numpy(r["f1"]).to(dev)
        slot = Slot(shp, 64, 4, 0.1, 0.5, 3, 1, 1.0, 1.0, dev)
        res = {}
        for b in ((1, 4) if args.pmc else (1, 2, 4)):
            reqs = [Request(j, shp, f0, f1) for j in range(b)]
            slot.launch(reqs)
            slot.stream.synchronize()
            slot.reqs = None
            if args.pmc:
                continue
            nat.timing_reset()
            nat.timing_enable(True)
            st, en = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            st.record(slot.stream)
            for _ in range(args.steps):
                slot.launch(reqs)
                slot.reqs = None
            en.record(slot.stream)
            slot.stream.synchronize()
            nat.timing_enable(False)
            kt = nat.timing_read()
            res[f"batch_{b}"] = {"ms_per_frame": round(st.elapsed_time(en) / args.steps / b, 4),
                                 "kernel_ms_per_frame": {k: round(v[0] / v[1] / b, 4) for k, v in kt.items()}}
        out["x".join(str(v) for v in shp)] = res
        del slot
        torch.cuda.empty_cache()
    if args.pmc:
        print(json.dumps({"pmc_run": True, "workload": "serving_batch", "shape": [384, 1248, 192],
                          "batches": [1, 4]}), flush=True)
        return {"pmc_run": True}
    line = common_line(args, "ms per frame by batch size (one serving slot)", out["384x1248x192"]["batch_4"]["ms_per_frame"],
                       "ms/frame", out["384x1248x192"]["batch_4"]["ms_per_frame"],
                       {"workload": args.workload, "feature_channels": 64, "iters": 1, "launches": args.steps},
                       higher=False)
    line["by_shape"] = out
    return line


def other_workloads(args, torch, benches):
    """Compact results of the semantic, flow and serving workloads for the default line (a few
    steps each; every workload also runs on its own with --workload)."""
    import copy
    sub = copy.copy(args)
    sub.steps, sub.warmup = 5, 2
    out = {}
    for name in ("semantic_fb_2x512x1024_c19", "semantic_pixel_fb_2x512x1024_c19", "flow_fb_256x512_r12",
                 "serving_c4"):
        sub.workload = name
        t0 = time.perf_counter()
        try:
            r = benches[name](sub, torch)
        except Exception as e:  # reported, never allowed to cost the headline number
            out[name] = {"error": repr(e)}
            continue
        if name == "serving_c4":
            c = {"frames_per_s": r["value"], "sequential_frames_per_s": r["sequential_frames_per_s"],
                 "speedup_vs_sequential": r["speedup_vs_sequential"],
                 "latency_ms_at_offered_load": r["latency_ms_at_offered_load"]}
        else:
            c = {"metric": r["metric"], "value": r["value"], "unit": r["unit"], "ms_per_step": r["ms_per_step"],
                 "steps": sub.steps, "config": r["config"], "timings_ms": r.get("timings_ms")}
            ri = r.get("roofline_iteration")
            if ri:
                c["roofline_iteration"] = {k: ri[k] for k in ("achieved", "peak", "unit", "frac", "alg_bytes_per_pl",
                                                              "iteration_ms")}
            if "sad" in r:
                c["sad"] = r["sad"]
        c["wall_s"] = round(time.perf_counter() - t0, 2)
        out[name] = c
        torch.cuda.empty_cache()
    return out


def measured_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc summary
    (FETCH_SIZE + WRITE_SIZE, calibrated with a copy kernel), if it covers this workload."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("workload") != workload or kernel not in t.get("kernels", {}):
        return None
    k = t["kernels"][kernel]
    return {"bytes_per_launch": k["hbm_bytes_per_launch"], "source": os.path.relpath(TRAFFIC_FILE, ROOT)}


def kernel_class(name):
    """Kernel class (as booked by libbpl_amd's timing) of a rocprofv3 kernel name."""
    import re
    base = name.split("(")[0] if "(anonymous" not in name else name.replace("(anonymous namespace)::", "").split("(")[0]
    if "copyBuffer" in name:
        return "calibration_copy"
    if "lbp_bwd_combine" in base or "lbp_bwd_pair_rows" in base:
        return "lbp_bwd_combine"
    if "lbp_bwd_reduce" in base:
        return "lbp_bwd_reduce"
    for stem, cls in (("sad_stereo_fwd", "sad_stereo_fwd"), ("sad_stereo_bwd", "sad_stereo_bwd"),
                      ("sad_stereo_softmax", "sad_stereo_softmax"), ("sad_flow_fwd", "sad_flow_fwd"),
                      ("sad_flow_bwd", "sad_flow_bwd"), ("softmax_labels", "softmax_labels"),
                      ("beliefs_k", "lbp_beliefs")):
        if stem in base:
            return cls
    if "lbp_fwd_stereo_pair" in base:  # the paired horizontal forward (base-plane launches)
        return "lbp_fwd_horizontal"
    m = re.search(r"lbp_(fwd|bwd)_(stereo_w64|semg|semp|phase)<([^>]*)>", base)
    if not m:
        return None
    targs = [t.strip() for t in m.group(3).split(",")]
    vert = targs[3] if m.group(2) == "phase" else targs[1]
    return f"lbp_{m.group(1)}_{'vertical' if vert == 'true' else 'horizontal'}"


# per-phase canonical bytes of the semantic workloads (per pixel: edge weights only, read in
# the forward, read and their gradient written in the backward)
PHASE_B_PL_SEMANTIC = {"lbp_fwd_horizontal": (22.0, 8.0), "lbp_fwd_vertical": (22.0, 8.0),
                       "lbp_bwd_vertical": (22.0, 16.0), "lbp_bwd_horizontal": (30.0, 16.0)}
# per-pixel matrices: each pass reads its direction's C x C matrix per image and pixel (4C B/pl) in the
# forward, and reads it and writes its gradient in the backward (8C B/pl), C = 19
PHASE_B_PL_SEMANTIC_PIXEL = {"lbp_fwd_horizontal": (22.0 + 2 * 76.0, 8.0), "lbp_fwd_vertical": (22.0 + 2 * 76.0, 8.0),
                             "lbp_bwd_vertical": (22.0 + 4 * 76.0, 16.0), "lbp_bwd_horizontal": (30.0 + 4 * 76.0, 16.0)}
PMC_WORKLOADS = {  # pixel*labels of one LBP iteration, labels, canonical phase bytes, output file
    "stereo_fb_384x1248_d192": (384 * 1248 * 192, 192, PHASE_B_PL, TRAFFIC_FILE),
    "semantic_fb_2x512x1024_c19": (2 * 512 * 1024 * 19, 19, PHASE_B_PL_SEMANTIC, None),
    "flow_fb_256x512_r12": (256 * 512 * 25, 25, PHASE_B_PL, None),  # one volume (u or v) per launch
    "semantic_pixel_fb_2x512x1024_c19": (2 * 512 * 1024 * 19, 19, PHASE_B_PL_SEMANTIC_PIXEL, None),
}


def pmc_summary(dirs, workload="stereo_fb_384x1248_d192"):
    """profiles/traffic.json from rocprofv3 --pmc runs of `bench.py --pmc`: per kernel class the
    HBM bytes per launch (FETCH_SIZE and WRITE_SIZE, in KiB per dispatch, each scaled by what it
    reads for the 256 MiB calibration copy) and the SQ counters collected."""
    import csv
    import glob
    import collections
    vals = collections.defaultdict(lambda: collections.defaultdict(list))
    for d in dirs:
        for fn in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            with open(fn) as f:
                for r in csv.DictReader(f):
                    cls = kernel_class(r["Kernel_Name"])
                    if cls:
                        vals[cls][r["Counter_Name"]].append(float(r["Counter_Value"]))
    mean = {c: {k: sum(v) / len(v) for k, v in d.items()} for c, d in vals.items()}
    # the 256 MiB copy is the largest copy dispatch (small host-device copies use the same kernel)
    cal = {k: max(v) for k, v in vals.get("calibration_copy", {}).items()}
    copy_bytes = 256 << 20
    fs = copy_bytes / (cal["FETCH_SIZE"] * 1024) if "FETCH_SIZE" in cal else None
    ws = copy_bytes / (cal["WRITE_SIZE"] * 1024) if "WRITE_SIZE" in cal else None
    pl, n_lab, phase_b, out_file = PMC_WORKLOADS[workload]
    out_file = out_file or os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    out = {"workload": workload,
           "command": "cd /tmp && export TMPDIR=/tmp && rocprofv3 --pmc <counters> --output-format csv "
                      f"-d $REPO/gpurun_out/pmc_<name> -o run -- python3 $REPO/bench.py --pmc --workload {workload} "
                      "(one run per counter set: FETCH_SIZE; WRITE_SIZE; SQ counters; TCC counters), then "
                      f"python3 bench.py --pmc-summary gpurun_out/pmc_* --workload {workload}",
           "calibration": {"kernel": "__amd_rocclr_copyBuffer (torch copy_ of 256 MiB fp32)",
                           "bytes_read": copy_bytes, "bytes_written": copy_bytes,
                           "FETCH_SIZE_KiB": cal.get("FETCH_SIZE"), "WRITE_SIZE_KiB": cal.get("WRITE_SIZE"),
                           "fetch_scale": round(fs, 4) if fs else None, "write_scale": round(ws, 4) if ws else None},
           "kernels": {}}
    canon = {k: v[0] + v[1] / n_lab for k, v in phase_b.items()}
    if workload == "stereo_fb_384x1248_d192":
        canon["lbp_bwd_horizontal"] += ACC_B_PL * 2 / 3  # 2 of the 3 chained calls accumulate
    else:
        canon["lbp_bwd_horizontal"] += ACC_B_PL / 2      # 1 of the 2 chained calls accumulates
    rows = []
    for cls in sorted(mean):
        if cls == "calibration_copy":
            continue
        m = mean[cls]
        k = {"launches_sampled": len(next(iter(vals[cls].values())))}
        if fs and ws and "FETCH_SIZE" in m and "WRITE_SIZE" in m:
            fb, wb = m["FETCH_SIZE"] * 1024 * fs, m["WRITE_SIZE"] * 1024 * ws
            k.update({"fetch_bytes_per_launch": int(fb), "write_bytes_per_launch": int(wb),
                      "hbm_bytes_per_launch": int(fb + wb)})
            if cls.startswith("lbp"):
                k["hbm_bytes_per_pl"] = round((fb + wb) / pl, 3)
                if cls in canon:
                    k["canonical_bytes_per_pl"] = round(canon[cls], 3)
        k["sq"] = {c: int(v) for c, v in m.items() if c.startswith("SQ_")}
        tcc = {c: int(v) for c, v in m.items() if c.startswith("TCC_")}
        if tcc:
            k["tcc"] = tcc
        out["kernels"][cls] = k
        rows.append((cls, k))
    with open(out_file, "w") as f:
        json.dump(out, f, indent=1)
    print(f"calibration: fetch x{out['calibration']['fetch_scale']}, write x{out['calibration']['write_scale']}")
    print("| kernel class | HBM bytes / launch | fetch | write | B/pl measured | B/pl canonical |")
    print("|---|---:|---:|---:|---:|---:|")
    for cls, k in rows:
        if "hbm_bytes_per_launch" in k:
            print(f"| {cls} | {k['hbm_bytes_per_launch'] / 1e9:.3f} GB | {k['fetch_bytes_per_launch'] / 1e9:.3f} GB | "
                  f"{k['write_bytes_per_launch'] / 1e9:.3f} GB | {k.get('hbm_bytes_per_pl', '')} | "
                  f"{k.get('canonical_bytes_per_pl', '')} |")
    print("| kernel class | waves | VALU / wave | SALU / wave | cycles / wave |")
    print("|---|---:|---:|---:|---:|")
    for cls, k in rows:
        q = k["sq"]
        if "SQ_WAVES" in q and q["SQ_WAVES"]:
            w = q["SQ_WAVES"]
            print(f"| {cls} | {w} | {q.get('SQ_INSTS_VALU', 0) / w:.0f} | {q.get('SQ_INSTS_SALU', 0) / w:.0f} | "
                  f"{q.get('SQ_WAVE_CYCLES', 0) / w:.0f} |")


def pmc_steps(step, torch, dev, info):
    """pmc_run for a step function: one warm-up step, one step, the 256 MiB calibration copy."""
    step()
    torch.cuda.synchronize()
    step()
    torch.cuda.synchronize()
    src = torch.ones(64 * 1024 * 1024, device=dev)  # 256 MiB of fp32
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"pmc_run": True, "steps": 2, **info, "calibration_copy_bytes": 2 * src.numel() * 4}), flush=True)
    return {"pmc_run": True}


def pmc_run(plan, torch):
    """One warm-up step, one step and a 256 MiB device-to-device copy, so that a
    rocprofv3 --pmc pass sees every kernel class of the step once per iteration and a
    copy kernel of known size (256 MiB read + 256 MiB written) to calibrate the counters."""
    plan.step()
    torch.cuda.synchronize()
    plan.step()
    torch.cuda.synchronize()
    src = torch.ones(64 * 1024 * 1024, device=plan.dev)  # 256 MiB of fp32
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"pmc_run": True, "steps": 2, "iters_per_step": plan.iters, "pl": plan.pl(),
                      "calibration_copy_bytes": 2 * src.numel() * 4}), flush=True)


def main():
    args = parse()
    if args.pmc_summary:
        pmc_summary(args.pmc_summary, args.workload)
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: this build runs on exactly one GPU; the multi-GPU path is not built "
              f"(--gpus {args.gpus}, WORLD_SIZE {world}).", file=sys.stderr)
        sys.exit(2)
    import torch
    from bplayers_amd import workloads

    others = {"semantic_fb_2x512x1024_c19": bench_semantic, "semantic_pixel_fb_2x512x1024_c19": bench_semantic_pixel,
              "flow_fb_256x512_r12": bench_flow,
              "serving_c4": bench_serving, "serving_batch": bench_serving_batch}
    if args.workload in others:
        print(json.dumps(others[args.workload](args, torch)), flush=True)
        return
    if args.workload == "stereo_fb_384x1248_d192":
        wl = workloads.stereo_full()
        backward = True
    else:
        wl = workloads.stereo_small()
        backward = False
    plan = StereoPlan(wl, wl["iters"], rows=args.rows, backward=backward)
    torch.cuda.synchronize()
    nat = plan.nat

    if args.pmc:
        pmc_run(plan, torch)
        return

    for _ in range(args.warmup):
        plan.step()
    torch.cuda.synchronize()
    # the timed region; the library brackets each of its launches with HIP events on the
    # launch stream (two event records per launch, no synchronisation)
    nat.timing_reset()
    nat.timing_enable(True)
    t0 = time.perf_counter()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(args.steps):
        plan.step()
    end.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    nat.timing_enable(False)
    kt = nat.timing_read()
    ms_step = start.elapsed_time(end) / args.steps

    pl = plan.pl()
    iters = plan.iters
    value = pl * iters / (ms_step * 1e-3) / 1e9
    C = plan.C
    avg = {k: v[0] / v[1] for k, v in kt.items()}   # ms per launch

    # roofline of the dominant kernel class and of the whole iteration
    lbp_k = {k: kt[k][0] for k in kt if k in PHASE_B_PL}
    dom = max(lbp_k, key=lbp_k.get)
    per_pl, per_px = PHASE_B_PL[dom]
    # the backward calls after the top one add to grad_cost (read + write): +4 B/pl on average
    # over the chain's calls (iters - 1 of iters)
    acc_pl = ACC_B_PL * (iters - 1) / iters if backward else 0.0
    if dom == "lbp_bwd_horizontal":
        per_pl += acc_pl
    dom_bytes = pl * (per_pl + per_px / C)
    dom_achieved = dom_bytes / (avg[dom] * 1e-3) / 1e9
    fwd_b = FWD_B_PL + (32.0 + 16.0) / C
    bwd_b = BWD_B_PL + (64.0 + 32.0) / C
    it_classes = LBP_CLASSES if backward else LBP_CLASSES[:2]
    it_ms = sum(avg[k] for k in it_classes if k in avg)
    it_bytes = pl * (fwd_b + (bwd_b + acc_pl if backward else -4.0))  # forward-only: no argmin writes
    it_achieved = it_bytes / (it_ms * 1e-3) / 1e9
    traffic = measured_traffic(args.workload, dom)
    res = {f"{k}_ms": v for k, v in avg.items()}

    # SAD cost volume (same workload), timed the same way
    cost_buf = torch.empty((plan.N, plan.H, plan.W, plan.max_disp - plan.min_disp + 1), device=plan.dev)
    d0 = torch.empty_like(plan.f0)
    d1 = torch.empty_like(plan.f1)
    for rep in range(6):
        if rep == 1:
            nat.timing_reset()
            nat.timing_enable(True)
        nat.call("bpl_sad_stereo_fwd", plan.sp(plan.f0), plan.sp(plan.f1), plan.sp(cost_buf), plan.N,
                 plan.f0.shape[1], plan.H, plan.W, plan.min_disp, plan.max_disp, 1.0, nat.stream_of(plan.f0))
        if backward:
            nat.call("bpl_sad_stereo_bwd", plan.sp(plan.f0), plan.sp(plan.f1), plan.sp(plan.sad_grad), plan.sp(d0),
                     plan.sp(d1), plan.N, plan.f0.shape[1], plan.H, plan.W, plan.min_disp, plan.max_disp,
                     nat.stream_of(plan.f0))
    torch.cuda.synchronize()
    nat.timing_enable(False)
    for k, v in nat.timing_read().items():
        res[f"{k}_ms"] = v[0] / v[1]
    t_sad = res["sad_stereo_fwd_ms"]

    line = {
        "metric": "G(pixel*label)/s per LBP iter fwd+bwd" if backward else "G(pixel*label)/s per LBP iter fwd",
        "value": round(value, 4),
        "unit": "G(pixel*label)/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": args.workload + (f"_crop{args.rows}" if args.rows else ""), "N": plan.N,
                   "H": plan.H, "W": plan.W, "labels": plan.C, "iters": iters, "delta": plan.delta,
                   "pairwise": "per-pixel per-direction P1/P2 (K=6 jump)" if backward else "global P1=0.1 P2=0.5",
                   "grad_mode": "exact" if backward else None},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(dom_achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(dom_achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic["bytes_per_launch"] if traffic else None,
                     "alg_bytes_per_launch": round(dom_bytes), "alg_bytes_per_pl": round(per_pl + per_px / C, 3),
                     "avg_ms": round(avg[dom], 4), "launches": kt[dom][1],
                     "traffic_source": traffic["source"] if traffic else None},
        "roofline_iteration": {"bound": "hbm", "kernels": list(it_classes), "achieved": round(it_achieved, 2),
                               "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(it_achieved / HBM_PEAK_GBS, 4),
                               "alg_bytes_per_pl": round(fwd_b + (bwd_b + acc_pl if backward else -4.0), 3),
                               "iteration_ms": round(it_ms, 4)},
        "timings_ms": {k: round(v, 4) for k, v in res.items()},
        "sad": {"fwd_G_pd_per_s": round(pl / (t_sad * 1e-3) / 1e9, 3),
                "bwd_G_pd_per_s": round(pl / (res["sad_stereo_bwd_ms"] * 1e-3) / 1e9, 3) if backward else None},
        "wall_s": round(wall, 3),
    }
    if args.workload == "stereo_fb_384x1248_d192" and not args.no_others and not args.rows:
        line["other_workloads"] = other_workloads(args, torch, others)
    if not args.no_cpu:
        try:
            line["cpu_baseline"] = cpu_baseline(wl, (plan.N, plan.H, plan.W, plan.C), args.cpu_seconds)
        except Exception as e:  # the baseline is reported, never required for the GPU number
            line["cpu_baseline"] = {"value": None, "error": repr(e)}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
