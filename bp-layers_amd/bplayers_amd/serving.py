"""Continuous-batching stereo serving on one GPU.

Requests (a stereo feature pair of one of several shape classes H x W x D) arrive over
time.  Each request is solved forward-only: cost = -softmax(-SAD/T) (SAD and softmax in one
kernel) -> ``iters`` LBP iterations with a global P1/P2 pairwise -> winner-take-all disparity
of the beliefs -- all on libbpl_amd kernels, with no allocation on the request path.

Design:
  * a *slot* owns the resident device buffers for a batch of up to ``max_batch``
    requests of one shape class (features, cost volume, messages, output) and
    its own HIP stream; slots are created lazily while the total resident bytes stay
    within the HBM budget (admission), and reused;
  * the scheduler loop admits requests as they arrive, groups waiting requests of one
    shape class into a batch (oldest first) whenever a slot of that class is free, and
    launches it asynchronously; batches of different slots overlap on the device;
  * completion is detected with per-slot events; latency = completion - arrival.
  * ``max_inflight`` (optional) caps the batches on the device at once: batches of
    different slots share the GPU's compute units, so every in-flight batch runs slower; the
    cap trades that overlap for shorter individual solves;
  * ``sequential=True`` is the baseline: one request at a time, batch 1, same kernels.

The kernels are batch-invariant (every batch element is solved by its own waves with a
fixed reduction order), so a request's result is bitwise the same alone and in a mix.
"""
import collections
import time

import numpy as np
import torch

from . import _native as nat
from .ops.lbp_stereo.message_passing_op_cuda import pack_affinities


class AdmissionError(RuntimeError):
    """A request can never be admitted under the configured HBM budget."""


class Request:
    __slots__ = ("rid", "shape", "f0", "f1", "arrival", "start", "done", "result")

    def __init__(self, rid, shape, f0, f1, arrival=0.0):
        self.rid = rid
        self.shape = shape  # (H, W, D)
        self.f0, self.f1 = f0, f1  # device tensors [1,C,H,W]
        self.arrival = arrival
        self.start = self.done = None
        self.result = None


def slot_bytes(shape, C, batch, iters=1):
    H, W, D = shape
    px = H * W
    # features, cost, messages, the forward's base-plane workspace, edges, disparities; jump
    return batch * (2 * C * px * 4 + px * D * 4 + 4 * px * D * 4 + px * D * 4 + 4 * px * 4 + px * 4) + 4 * 6 * px * 4


class Slot:
    def __init__(self, shape, C, max_batch, P1, P2, delta, iters, T_match, T_bel, device):
        H, W, D = shape
        self.shape, self.C, self.max_batch = shape, C, max_batch
        self.iters, self.delta, self.T_match, self.T_bel = iters, delta, T_match, T_bel
        B = max_batch
        dev = device
        self.stream = torch.cuda.Stream(device=dev)  # default priority
        self.f0 = torch.empty((B, C, H, W), device=dev)
        self.f1 = torch.empty((B, C, H, W), device=dev)
        self.cost = torch.empty((B, H, W, D), device=dev)
        self.msgs = torch.empty((B, 4, H, W, D), device=dev)
        self.base = torch.empty((B, H, W, D), device=dev)  # bpl_lbp_fwd_ws workspace
        self.edge = torch.ones((B, 4, H, W), device=dev)
        self.disp = torch.empty((B, H, W), device=dev)
        aff = torch.zeros((1, 4, 3, H, W), device=dev)
        aff[:, :, :2] = P1
        aff[:, :, 2] = P2
        self.jump = pack_affinities(aff, torch.zeros((1, 4, H, W), device=dev))
        self.event = torch.cuda.Event()
        self.reqs = None
        self.bytes = slot_bytes(shape, C, B)
        torch.cuda.synchronize(dev)

    def launch(self, reqs):
        n = len(reqs)
        H, W, D = self.shape
        C = self.C
        p = nat.ptr
        # the slot's stream first waits for the caller's stream: that orders this batch after
        # the copy of the previous batch's results (finish() clones them on the caller's
        # stream) and after whatever produced the requests' features
        self.stream.wait_stream(torch.cuda.current_stream(self.stream.device))
        with torch.cuda.stream(self.stream):
            st = ctypes_stream(self.stream)
            for i, r in enumerate(reqs):
                self.f0[i].copy_(r.f0[0], non_blocking=True)
                self.f1[i].copy_(r.f1[0], non_blocking=True)
            nat.call("bpl_sad_stereo_softmax_fwd", p(self.f0), p(self.f1), p(self.cost), n, C, H, W, 0, D - 1,
                     -1.0 / self.T_match, -1.0, st)
            self.msgs[:n].zero_()
            for it in range(self.iters):
                nat.call("bpl_lbp_fwd_ws", nat.LBP_STEREO, p(self.cost), p(self.jump), 6, 1, p(self.edge),
                         p(self.msgs), p(self.msgs), None, None, None, n, H, W, D, self.delta, p(self.base),
                         self.base.numel() * 4, nat.FWD_MSG_IN_ZERO if it == 0 else 0, st)
            nat.call("bpl_lbp_beliefs", p(self.cost), p(self.msgs), None, p(self.disp), n, H, W, D, self.T_bel, st)
            self.event.record(self.stream)
        self.reqs = reqs

    def poll(self):
        return self.reqs is not None and self.event.query()

    def finish(self, keep_results):
        """Called once the slot's event has completed; the results are copied on the caller's
        current stream (the next launch() of this slot waits for that stream)."""
        reqs = self.reqs
        if keep_results:
            for i, r in enumerate(reqs):
                r.result = self.disp[i].clone()
        self.reqs = None
        return reqs


def ctypes_stream(s):
    import ctypes
    return ctypes.c_void_p(s.cuda_stream)


class StereoServer:
    def __init__(self, device="cuda:0", C=64, hbm_budget=32 << 30, max_batch=4, max_slots_per_class=2, iters=1,
                 P1=0.1, P2=0.5, delta=3, T_match=1.0, T_bel=1.0, sequential=False, max_inflight=None):
        self.device = torch.device(device)
        self.C = C
        self.budget = int(hbm_budget)
        self.max_batch = 1 if sequential else max_batch
        self.max_slots = 1 if sequential else max_slots_per_class
        self.sequential = sequential
        self.max_inflight = 1 if sequential else max_inflight
        self.params = dict(P1=P1, P2=P2, delta=delta, iters=iters, T_match=T_match, T_bel=T_bel)
        self.slots = collections.defaultdict(list)
        self.resident = 0

    def _slot_for(self, shape):
        free = [s for s in self.slots[shape] if s.reqs is None]
        if free:
            return free[0]
        if len(self.slots[shape]) >= self.max_slots:
            return None
        need = slot_bytes(shape, self.C, self.max_batch)
        if need > self.budget:
            raise AdmissionError(f"a {shape} slot needs {need} bytes, above the HBM budget of {self.budget}")
        if self.resident + need > self.budget:
            # reclaim an idle slot of another class
            for shp, lst in self.slots.items():
                for s in lst:
                    if s.reqs is None and shp != shape:
                        lst.remove(s)
                        self.resident -= s.bytes
                        del s
                        torch.cuda.empty_cache()
                        if self.resident + need <= self.budget:
                            break
                if self.resident + need <= self.budget:
                    break
            if self.resident + need > self.budget:
                return None
        p = self.params
        s = Slot(shape, self.C, self.max_batch, p["P1"], p["P2"], p["delta"], p["iters"], p["T_match"], p["T_bel"],
                 self.device)
        self.slots[shape].append(s)
        self.resident += s.bytes
        return s

    def warm(self, shapes):
        """Create one slot per shape class ahead of time (not on the request path)."""
        for shp in shapes:
            self._slot_for(shp)

    def serve(self, requests, keep_results=False):
        """Serve ``requests`` (sorted by arrival, seconds from the start).  Returns stats."""
        for r in requests:
            need = slot_bytes(r.shape, self.C, self.max_batch)
            if need > self.budget:
                raise AdmissionError(f"request {r.rid} of shape {r.shape} needs {need} bytes, above the budget")
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        queues = collections.OrderedDict()
        inflight = []
        done = []
        torch.cuda.synchronize(self.device)
        t0 = time.perf_counter()
        while pending or any(queues.values()) or inflight:
            now = time.perf_counter() - t0
            while pending and pending[0].arrival <= now:
                r = pending.popleft()
                queues.setdefault(r.shape, collections.deque()).append(r)
            for s in list(inflight):
                if s.poll():
                    t = time.perf_counter() - t0
                    for r in s.finish(keep_results):
                        r.done = t
                        done.append(r)
                    inflight.remove(s)
            launched = False
            cap = self.max_inflight
            if cap is None or len(inflight) < cap:
                # oldest waiting request first decides the next shape class
                order = sorted((q[0].arrival, shp) for shp, q in queues.items() if q)
                for _, shp in order:
                    if cap is not None and len(inflight) >= cap:
                        break
                    s = self._slot_for(shp)
                    if s is None:
                        continue
                    q = queues[shp]
                    batch = [q.popleft() for _ in range(min(len(q), s.max_batch))]
                    t = time.perf_counter() - t0
                    for r in batch:
                        r.start = t
                    s.launch(batch)
                    inflight.append(s)
                    launched = True
            if not launched and not pending and not any(queues.values()) and inflight:
                time.sleep(20e-6)
            elif not launched and pending and not any(queues.values()) and not inflight:
                wait = pending[0].arrival - (time.perf_counter() - t0)
                if wait > 0:
                    time.sleep(min(wait, 1e-3))
        torch.cuda.synchronize(self.device)
        lat = np.array([r.done - r.arrival for r in done])
        span = max(r.done for r in done) - min(r.arrival for r in done)
        return {"frames": len(done), "frames_per_s": len(done) / span, "p50_ms": float(np.percentile(lat, 50) * 1e3),
                "p99_ms": float(np.percentile(lat, 99) * 1e3), "mean_ms": float(lat.mean() * 1e3),
                "span_s": span, "resident_bytes": self.resident}


def make_requests(pool, n, rate_hz, seed=0):
    """``n`` requests with Poisson arrivals at ``rate_hz``, shape classes drawn uniformly;
    ``pool`` maps shape -> list of (f0, f1) device tensors."""
    rng = np.random.default_rng(seed)
    shapes = sorted(pool.keys())
    t = 0.0
    out = []
    for i in range(n):
        t += rng.exponential(1.0 / rate_hz)
        shp = shapes[rng.integers(len(shapes))]
        f0, f1 = pool[shp][rng.integers(len(pool[shp]))]
        out.append(Request(i, shp, f0, f1, t))
    return out
