"""Serving loop: continuous batching of projector and SIRT requests.

Requests (forward A, backprojection AT, or an n-iteration SIRT
reconstruction) arrive over time.  The server keeps every request's data
resident on the device and admits a request when its buffers fit the memory
budget.  It then runs *ticks*.  In one tick every admitted request advances
by one step:

  1. one persistent batched A launch over all requests that need an A now
     (plain A requests, SIRT row sums A1, SIRT A x);
  2. SIRT residual update r = R (b - A x);
  3. one persistent batched AT launch over all requests that need an AT
     (plain AT requests, SIRT column sums AT1, SIRT AT r);
  4. SIRT update x += C AT r.

Requests that arrive while a tick is on the GPU join the next tick, so work
of different requests overlaps in the same launches ("continuous batching").
An instance whose own launch fills the device runs as a plain single launch;
the others run with the kernel variant chosen from their own shape (one
batched launch per variant group).  The choice depends only on the
instance's shape, and both paths add the same terms in the same order, so a
request's result does not depend on what else is in flight: alone and in
the mix it is bitwise identical.

Latency = completion - arrival; throughput = voxel-angles / makespan, where
an A or AT of an (SY, SX, H) volume at n_angles counts SY*SX*H*n_angles and a
SIRT request counts (2 + 2 n_iter) of them.

No counterpart in the reference (a library of two operators).
"""
import math
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib, ops, recon
from .batch import BatchPlan, default_grid

KINDS = ("A", "AT", "SIRT")


@dataclass
class Request:
    rid: int
    kind: str                  # "A", "AT" or "SIRT"
    arrival: float             # seconds from the start of the run
    data: torch.Tensor         # A: volume (SY,SX,H); AT: sinogram (A,PX,H); SIRT: sinogram b
    angles: torch.Tensor
    SY: int
    SX: int
    n_iter: int = 10
    # filled in by the server
    result: torch.Tensor = None
    t_admit: float = None
    t_done: float = None
    state: dict = field(default_factory=dict)

    @property
    def H(self):
        return int(self.data.shape[2])

    @property
    def PX(self):
        return int(self.data.shape[1]) if self.kind != "A" else max(self.SY, self.SX)

    @property
    def n_angles(self):
        return int(self.angles.numel())

    def unit_work(self):
        return self.SY * self.SX * self.H * self.n_angles

    def work(self):
        return self.unit_work() * ((2 + 2 * self.n_iter) if self.kind == "SIRT" else 1)

    def bytes_needed(self):
        v = self.SY * self.SX * self.H * 4
        s = self.n_angles * self.PX * self.H * 4
        if self.kind == "A":
            return v + s
        if self.kind == "AT":
            return s + v
        return 4 * v + 4 * s  # b, R, tmp_s, ones_s / x, C, tmp_v, ones_v


def _safe_reciprocal_(t, eps=1e-6):
    """t <- 1/t where t > eps, else 0 (the SIRT row / column weights)."""
    mask = t > eps
    t.clamp_min_(eps).reciprocal_().mul_(mask)
    return t


def _variant(op, H):
    return int(_lib.load().vp_default_variant(0 if op == "A" else 1, H))


class Server:
    """Continuous-batching server on one device."""

    def __init__(self, device="cuda", mem_budget=None, max_active=256, grid=None):
        self.device = torch.device(device)
        if mem_budget is None:
            free, _total = torch.cuda.mem_get_info(self.device)
            mem_budget = int(0.8 * free)
        self.mem_budget = int(mem_budget)
        self.max_active = max_active
        self.grid = default_grid(self.device) if grid is None else grid
        self.cus = torch.cuda.get_device_properties(self.device).multi_processor_count
        self.active = []
        self.in_use = 0
        self.ticks = 0
        self._ev_end = None  # event after the last launch of the latest tick

    # ------------------------------------------------------------ admission
    def can_admit(self, req):
        return (len(self.active) < self.max_active
                and self.in_use + req.bytes_needed() <= self.mem_budget)

    def admit(self, req, now):
        req.t_admit = now
        self.in_use += req.bytes_needed()
        st = req.state
        dev = self.device
        if req.kind == "A":
            req.result = torch.empty((req.n_angles, req.PX, req.H), device=dev)
        elif req.kind == "AT":
            req.result = torch.empty((req.SY, req.SX, req.H), device=dev)
        else:
            vshape = (req.SY, req.SX, req.H)
            sshape = tuple(req.data.shape)
            st["x"] = torch.zeros(vshape, device=dev)
            st["R"] = torch.empty(sshape, device=dev)
            st["C"] = torch.empty(vshape, device=dev)
            st["ts"] = torch.empty(sshape, device=dev)
            st["tv"] = torch.empty(vshape, device=dev)
            st["ones_v"] = torch.ones(vshape, device=dev)
            st["ones_s"] = torch.ones(sshape, device=dev)
            st["it"] = -1  # -1: weights not computed yet
            req.result = st["x"]
        self.active.append(req)

    # ------------------------------------------------------------ one tick
    def _fills_device(self, op, inp, out, angles, H):
        """Whether the instance's own launch has at least one workgroup per CU:
        then it runs as a plain single launch (batching adds nothing to it)."""
        if op == "A":
            (SY, SX, _), PX = inp.shape, out.shape[1]
        else:
            (SY, SX, _), PX = out.shape, inp.shape[1]
        code = 0 if op == "A" else 1
        nb = _lib.load().vp_batch_blocks(code, _variant(op, H), angles.numel(), SY, SX, H, PX)
        return nb >= self.cus

    def _launch(self, op, jobs):
        """jobs: list of (in, out, angles, H).  Instances that fill the device
        on their own run as plain single launches; the rest go into one
        batched launch per variant.  Both paths add the same terms in the same
        order, so a result does not depend on which path or batch it took."""
        groups = {}
        for j in jobs:
            if self._fills_device(op, *j):
                inp, out, angles, _H = j
                out.zero_()
                if op == "A":
                    ops.project(inp, angles, out.shape[1], out=out)
                else:
                    ops.backproject(inp, angles, out.shape[0], out.shape[1], out=out)
            else:
                groups.setdefault(_variant(op, j[3]), []).append(j)
        for v, js in sorted(groups.items()):
            BatchPlan(op, [j[0] for j in js], [j[1] for j in js], [j[2] for j in js],
                      accumulate=False, variant=v, grid=self.grid).launch()

    def tick(self):
        """Advance every admitted request by one step.  Single-operator
        requests are launched first (A, then AT), so they complete before the
        SIRT steps of the same tick; returns [(request, completion event)]."""
        a_plain, at_plain, a_sirt, at_sirt = [], [], [], []
        for r in self.active:
            st = r.state
            if r.kind == "A":
                a_plain.append((r.data, r.result, r.angles, r.H))
            elif r.kind == "AT":
                at_plain.append((r.data, r.result, r.angles, r.H))
            elif st["it"] < 0:
                a_sirt.append((st["ones_v"], st["R"], r.angles, r.H))
                at_sirt.append((st["ones_s"], st["C"], r.angles, r.H))
            else:
                a_sirt.append((st["x"], st["ts"], r.angles, r.H))
        ev_plain = None
        if a_plain:
            self._launch("A", a_plain)
        if at_plain:
            self._launch("AT", at_plain)
        if a_plain or at_plain:
            ev_plain = torch.cuda.Event(enable_timing=True)
            ev_plain.record()
        if a_sirt:
            self._launch("A", a_sirt)
        for r in self.active:
            st = r.state
            if r.kind == "SIRT":
                if st["it"] < 0:
                    _safe_reciprocal_(st["R"])
                else:
                    torch.sub(r.data, st["ts"], out=st["ts"])
                    st["ts"].mul_(st["R"])
                    at_sirt.append((st["ts"], st["tv"], r.angles, r.H))
        if at_sirt:
            self._launch("AT", at_sirt)
        done = []
        for r in self.active:
            st = r.state
            if r.kind != "SIRT":
                done.append(r)
            elif st["it"] < 0:
                _safe_reciprocal_(st["C"])
                st["it"] = 0
            else:
                st["x"].addcmul_(st["C"], st["tv"])
                st["it"] += 1
                if st["it"] >= r.n_iter:
                    done.append(r)
        ev_end = torch.cuda.Event(enable_timing=True)
        ev_end.record()
        self._ev_end = ev_end
        self.ticks += 1
        return [(r, ev_plain if r.kind != "SIRT" else ev_end) for r in done]

    def _retire(self, r, now):
        r.t_done = now
        self.in_use -= r.bytes_needed()
        for k in ("R", "C", "ts", "tv", "ones_v", "ones_s"):
            r.state.pop(k, None)
        self.active.remove(r)

    # ------------------------------------------------------------ run
    def run(self, requests):
        """Serve `requests` (sorted by arrival) in real time; returns them."""
        pending = sorted(requests, key=lambda r: r.arrival)
        for r in pending:
            if r.bytes_needed() > self.mem_budget:
                raise RuntimeError(f"request {r.rid} needs {r.bytes_needed()} B, over the budget")
        torch.cuda.synchronize(self.device)
        t0 = time.perf_counter()
        i = 0
        while i < len(pending) or self.active:
            now = time.perf_counter() - t0
            while i < len(pending) and pending[i].arrival <= now and self.can_admit(pending[i]):
                self.admit(pending[i], now)
                i += 1
            if not self.active:
                # idle: wait for the next arrival
                time.sleep(max(0.0, pending[i].arrival - now))
                continue
            done = self.tick()
            torch.cuda.synchronize(self.device)
            t_sync = time.perf_counter() - t0
            # completion on the host clock: the tick's end is observed at
            # t_sync, and the request's last launch finished earlier by the
            # GPU time between its event and the tick's end event (never
            # before the tick was launched, so never before admission)
            for r, ev in done:
                self._retire(r, t_sync - ev.elapsed_time(self._ev_end) * 1e-3)
        self.makespan = time.perf_counter() - t0
        return requests


def run_alone(req, device="cuda"):
    """The same request through the same kernels with nothing else in flight."""
    s = Server(device=device)
    clone = Request(req.rid, req.kind, 0.0, req.data, req.angles, req.SY, req.SX, req.n_iter)
    s.run([clone])
    return clone.result


def _run_one_plain(r):
    if r.kind == "A":
        return ops.project(r.data, r.angles, r.PX)
    if r.kind == "AT":
        return ops.backproject(r.data, r.angles, r.SY, r.SX)
    solver = recon.SIRT(r.angles, r.SY, r.SX, r.H, r.PX, device=r.data.device)
    x = torch.zeros((r.SY, r.SX, r.H), device=r.data.device)
    solver.iterate(x, r.data, r.n_iter)
    return x


def run_plain(requests):
    """One request at a time through the plain single-instance operators,
    back to back from time 0 (arrival times ignored); returns (results,
    seconds)."""
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = [_run_one_plain(r) for r in requests]
    torch.cuda.synchronize()
    return out, time.perf_counter() - t0


def run_plain_fifo(requests):
    """A server without batching: requests in arrival order, each run alone
    through the plain operators once it has arrived and the previous one is
    done.  Returns {"p50_ms", "p99_ms", "makespan_s"} (latency = completion -
    arrival), the baseline for the continuous-batching latencies."""
    pending = sorted(requests, key=lambda r: r.arrival)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    lat = []
    for r in pending:
        now = time.perf_counter() - t0
        if r.arrival > now:
            time.sleep(r.arrival - now)
        _run_one_plain(r)
        torch.cuda.synchronize()
        lat.append((time.perf_counter() - t0 - r.arrival) * 1e3)
    makespan = time.perf_counter() - t0
    lat = np.array(lat)
    return {"p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "makespan_s": makespan}


def mixed_workload(n, rate=200.0, seed=0, sizes=(128, 256, 512), heights=(64, 128, 256, 512),
                   n_angles=(90, 192, 360, 720, 1440), n_iter=10, device="cuda"):
    """Poisson request stream: 40% A, 40% AT, 20% SIRT; ellipsoid-phantom or
    N(0,1) inputs generated on the device (seeded)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    reqs = []
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        kind = rng.choice(KINDS, p=[0.4, 0.4, 0.2])
        S = int(rng.choice(sizes))
        H = int(rng.choice(heights))
        A = int(rng.choice(n_angles))
        ang = torch.tensor(np.linspace(0, np.pi, A, endpoint=False).astype(np.float32), device=device)
        phantom = rng.random() < 0.5
        g = torch.Generator(device=device).manual_seed(seed * 100003 + i)
        if kind == "A":
            data = (recon.ellipsoid_phantom(S, S, H, seed=seed * 100003 + i, device=device)
                    if phantom else torch.randn((S, S, H), device=device, generator=g))
        elif kind == "AT":
            data = torch.randn((A, S, H), device=device, generator=g)
        else:
            x_true = recon.ellipsoid_phantom(S, S, H, seed=seed * 100003 + i, device=device)
            data = ops.project(x_true, ang, S)
        reqs.append(Request(i, str(kind), t, data, ang, S, S, n_iter))
    return reqs


def latency_stats(requests):
    lat = np.array([(r.t_done - r.arrival) * 1e3 for r in requests])
    return {"p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean()), "max_ms": float(lat.max())}


def throughput(requests, seconds):
    return sum(r.work() for r in requests) / seconds / 1e9  # Gvoxel*angle/s
