"""Serving loop: continuous batching of projector and SIRT requests.

Requests (forward A, backprojection AT, or an n-iteration SIRT
reconstruction) arrive over time.  The server admits a request as soon as a
slot is free and its device buffers fit the memory budget, keeps its data
resident on the device, and never waits for the device between launches:

* a request whose own A and AT launches fill the device (at least one
  workgroup per CU) is enqueued as a whole -- its single launch, or its SIRT
  chain of 2 + 2 n_iter launches and elementwise steps -- on one of a few
  request streams, round-robin.  Requests on different streams run
  concurrently, so a request that arrives while another is on the GPU starts
  at once, on the CUs that the other's workgroups release;
* the other requests are batched: on the batching stream, each *tick* is
  one persistent batched A launch over every such request that needs an A
  now (plain A requests, SIRT row sums A1, SIRT A x), the SIRT residual
  updates, one batched AT launch (plain AT requests, SIRT column sums AT1,
  SIRT AT r) and the SIRT updates; every batched request advances one step
  per tick.  Requests that arrive while ticks are in flight join the next
  tick; up to `depth` ticks are queued on the device at a time, and the
  launch descriptors of an unchanged set of requests are reused.

The host loop admits, launches and polls GPU events (Event.query); it never
synchronizes.  A request's completion time is read off the GPU clock: the
elapsed time from an event recorded at the start of the run to the event
after the request's last launch.

Which path a request takes depends only on its own shape, and both paths
add the same terms in the same order (the batched kernels equal the single
launches bit for bit), so a request's result does not depend on what else is
in flight: alone and in the mix it is bitwise identical, and equal to the
plain operators' result.

Latency = completion - arrival; throughput = voxel-angles / makespan, where
an A or AT of an (SY, SX, H) volume at n_angles counts SY*SX*H*n_angles and a
SIRT request counts (2 + 2 n_iter) of them.

No counterpart in the reference (a library of two operators).
"""
import time
from collections import OrderedDict
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

    def __post_init__(self):
        if self.kind not in KINDS:
            raise ValueError(f"request kind must be one of {KINDS}, got {self.kind!r}")
        if self.kind == "SIRT" and int(self.n_iter) < 0:
            raise ValueError("a SIRT request needs n_iter >= 0")

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
    """t <- 1/t where t > eps, else 0 (the SIRT row / column weights; the
    same values as recon.SIRT's)."""
    mask = t > eps
    t.clamp_min_(eps).reciprocal_().mul_(mask)
    return t


def _variant(op, H):
    return int(_lib.load().vp_default_variant(0 if op == "A" else 1, H))


def _run_chain(r):
    """The plain single-instance operators for one request (on the current
    stream); returns its result."""
    if r.kind == "A":
        return ops.project(r.data, r.angles, r.PX)
    if r.kind == "AT":
        return ops.backproject(r.data, r.angles, r.SY, r.SX)
    solver = recon.SIRT(r.angles, r.SY, r.SX, r.H, r.PX, device=r.data.device)
    x = torch.zeros((r.SY, r.SX, r.H), device=r.data.device)
    solver.iterate(x, r.data, r.n_iter)
    return x


class Server:
    """Continuous-batching server on one device.

    mem_budget: bytes of request buffers admitted at a time (default 80 % of
    the free device memory); max_active: request slots; streams: request
    streams for requests that fill the device; depth: batched ticks queued
    on the device at a time."""

    def __init__(self, device="cuda", mem_budget=None, max_active=64, grid=None, streams=2,
                 depth=2):
        self.device = torch.device(device)
        if mem_budget is None:
            free, _total = torch.cuda.mem_get_info(self.device)
            mem_budget = int(0.8 * free)
        self.mem_budget = int(mem_budget)
        self.max_active = max_active
        self.grid = default_grid(self.device) if grid is None else grid
        self.cus = torch.cuda.get_device_properties(self.device).multi_processor_count
        self.depth = max(1, int(depth))
        # default-priority streams: one per concurrently running large request, one for ticks
        self.req_streams = [torch.cuda.Stream(device=self.device) for _ in range(max(1, streams))]
        self.tick_stream = torch.cuda.Stream(device=self.device)
        self._next_stream = 0
        self.active = []        # admitted, not retired
        self.batched = []       # the active requests served by ticks
        self.in_use = 0
        self.ticks = 0
        self._flight = []       # (event, [requests completing at it])
        self._ticks_in_flight = []  # tick-end events still running
        self._plans = OrderedDict()  # cached launch descriptors
        self._ev0 = None

    # ------------------------------------------------------------ admission
    def can_admit(self, req):
        return (len(self.active) < self.max_active
                and self.in_use + req.bytes_needed() <= self.mem_budget)

    def _blocks(self, op, r):
        code = 0 if op == "A" else 1
        return _lib.load().vp_batch_blocks(code, _variant(op, r.H), r.n_angles, r.SY, r.SX,
                                           r.H, r.PX)

    def fills_device(self, req):
        """Whether the request's own A and AT launches each have at least one
        workgroup per CU (then batching adds nothing to them)."""
        ops_needed = {"A": ("A",), "AT": ("AT",), "SIRT": ("A", "AT")}[req.kind]
        return all(self._blocks(op, req) >= self.cus for op in ops_needed)

    def admit(self, req, now):
        req.t_admit = now
        self.in_use += req.bytes_needed()
        self.active.append(req)
        main = torch.cuda.current_stream(self.device)
        if self.fills_device(req):
            st = self.req_streams[self._next_stream % len(self.req_streams)]
            self._next_stream += 1
            st.wait_stream(main)  # the request's data was written on the main stream
            with torch.cuda.stream(st):
                req.result = _run_chain(req)
                ev = torch.cuda.Event(enable_timing=True)
                ev.record(st)
            self._flight.append((ev, [req]))
            return
        self.tick_stream.wait_stream(main)
        st = req.state
        dev = self.device
        with torch.cuda.stream(self.tick_stream):
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
        self.batched.append(req)

    # ------------------------------------------------------------ one tick
    def _plan(self, op, v, jobs):
        """The batched launch of `jobs` [(in, out, angles)] under variant v,
        built once per set of buffers and reused while that set recurs."""
        key = (op, v) + tuple((j[0].data_ptr(), j[1].data_ptr(), j[2].data_ptr()) for j in jobs)
        plan = self._plans.get(key)
        if plan is None:
            plan = BatchPlan(op, [j[0] for j in jobs], [j[1] for j in jobs], [j[2] for j in jobs],
                             accumulate=False, variant=v, grid=self.grid)
            self._plans[key] = plan
            while len(self._plans) > 64:
                self._plans.popitem(last=False)
        else:
            self._plans.move_to_end(key)
        return plan

    def _launch(self, op, jobs):
        """jobs: [(in, out, angles, H)]: one batched launch per kernel variant."""
        groups = {}
        for j in jobs:
            groups.setdefault(_variant(op, j[3]), []).append(j)
        for v, js in sorted(groups.items()):
            self._plan(op, v, js).launch()

    def tick(self):
        """Advance every batched request by one step, on the batching stream,
        without waiting for the device.  Single-operator requests are launched
        first (A, then AT), so they complete before the SIRT steps of the same
        tick; returns [(request, completion event)]."""
        a_plain, at_plain, a_sirt, at_sirt = [], [], [], []
        for r in self.batched:
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
        done = []
        with torch.cuda.stream(self.tick_stream):
            ev_plain = None
            if a_plain:
                self._launch("A", a_plain)
            if at_plain:
                self._launch("AT", at_plain)
            if a_plain or at_plain:
                ev_plain = torch.cuda.Event(enable_timing=True)
                ev_plain.record(self.tick_stream)
            if a_sirt:
                self._launch("A", a_sirt)
            for r in self.batched:
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
            for r in self.batched:
                st = r.state
                if r.kind != "SIRT":
                    done.append(r)
                elif st["it"] < 0:
                    _safe_reciprocal_(st["C"])
                    st["it"] = 0
                    if r.n_iter <= 0:
                        done.append(r)
                else:
                    st["x"].addcmul_(st["C"], st["tv"])
                    st["it"] += 1
                    if st["it"] >= r.n_iter:
                        done.append(r)
            ev_end = torch.cuda.Event(enable_timing=True)
            ev_end.record(self.tick_stream)
        self.ticks += 1
        for r in done:
            self.batched.remove(r)
        plain = [r for r in done if r.kind != "SIRT"]
        sirt = [r for r in done if r.kind == "SIRT"]
        if plain:
            self._flight.append((ev_plain, plain))
        if sirt:
            self._flight.append((ev_end, sirt))
        self._ticks_in_flight.append(ev_end)
        return [(r, ev_plain if r.kind != "SIRT" else ev_end) for r in done]

    def _retire(self, r, t_done):
        r.t_done = t_done
        self.in_use -= r.bytes_needed()
        for k in ("R", "C", "ts", "tv", "ones_v", "ones_s"):
            r.state.pop(k, None)
        self.active.remove(r)

    def _poll(self):
        """Retire the requests whose completion events have fired; returns
        whether any did."""
        fired = False
        keep = []
        for ev, reqs in self._flight:
            if ev.query():
                t = self._ev0.elapsed_time(ev) * 1e-3
                for r in reqs:
                    self._retire(r, max(t, r.t_admit))
                fired = True
            else:
                keep.append((ev, reqs))
        self._flight = keep
        self._ticks_in_flight = [e for e in self._ticks_in_flight if not e.query()]
        return fired

    # ------------------------------------------------------------ run
    def run(self, requests):
        """Serve `requests` (sorted by arrival) in real time; returns them."""
        pending = sorted(requests, key=lambda r: r.arrival)
        for r in pending:
            if r.bytes_needed() > self.mem_budget:
                raise RuntimeError(f"request {r.rid} needs {r.bytes_needed()} B, over the budget")
        main = torch.cuda.current_stream(self.device)
        torch.cuda.synchronize(self.device)
        self._ev0 = torch.cuda.Event(enable_timing=True)
        self._ev0.record(main)
        t0 = time.perf_counter()
        i = 0
        while i < len(pending) or self.active:
            now = time.perf_counter() - t0
            while i < len(pending) and pending[i].arrival <= now and self.can_admit(pending[i]):
                self.admit(pending[i], now)
                i += 1
            if self.batched and len(self._ticks_in_flight) < self.depth:
                self.tick()
                continue
            if self._poll():
                continue
            if not self.active:
                # idle: wait for the next arrival
                time.sleep(max(0.0, pending[i].arrival - now))
                continue
            # work is in flight: look again shortly (new arrivals, completions)
            time.sleep(5e-5)
        torch.cuda.synchronize(self.device)
        for st in self.req_streams + [self.tick_stream]:
            main.wait_stream(st)
        self.makespan = time.perf_counter() - t0
        return requests


def run_alone(req, device="cuda"):
    """The same request through the same kernels with nothing else in flight."""
    s = Server(device=device)
    clone = Request(req.rid, req.kind, 0.0, req.data, req.angles, req.SY, req.SX, req.n_iter)
    s.run([clone])
    return clone.result


def _run_one_plain(r):
    return _run_chain(r)


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
