"""Serving loop: continuous batching of projector and SIRT requests.

Requests (forward A, backprojection AT, or an n-iteration SIRT
reconstruction) arrive over time.  The server admits a request as soon as a
slot is free and its device buffers fit the memory budget, keeps its data
resident on the device, and splits its work into *steps*: one projector
launch each (with the SIRT elementwise updates that follow it), so a plain
request is one step and an n-iteration SIRT request 2 + 2 n steps.

Steps go to device streams, each queued ahead of the device up to about
`queue_ms` of estimated GPU time (at least 2 and at most `depth` steps, so
the next step is already queued when one ends), and the host never waits
for the device: it admits new arrivals, polls GPU events for completions
and, whenever a stream's queue has room, submits the next step of that
stream's request with the least remaining work (shortest remaining work
first, ties in arrival order).  A request whose launches fill the device (at
least one workgroup per CU) is pinned to one of two request streams, the one
with less remaining work; the others share a third stream, where the next
steps of all of them with the same operator run as one persistent batched
launch (one per kernel variant; the launch descriptors of a recurring set of
requests are reused; a lone small request runs as a single launch in its
own tuned configuration).  Kernels of different streams run concurrently and
fill each other's idle CUs; a request that arrives while others are on the
GPU starts as soon as the step in flight on its stream ends, and a short
request overtakes a long SIRT chain ("continuous batching").
A request's completion time is read off the GPU
clock: the elapsed time from an event recorded at the start of the run to
the event after its last step.

Requests come in either layout (z_order 0: volumes (SY, SX, H), sinograms
(A, PX, H); z_order 2: (H, SY, SX) and (H, A, PX)); only requests of the
same layout are batched together.

Which path a request takes depends only on its own shape, and the batched
kernels add every output's terms in the same order as the single launches,
so a request's result does not depend on what else is in flight: alone and
in the mix it is bitwise identical, and equal to the plain operators'.

Latency = completion - arrival; throughput = voxel-angles / makespan, where
an A or AT of an (SY, SX, H) volume at n_angles counts SY*SX*H*n_angles and a
SIRT request counts (2 + 2 n_iter) of them.

No counterpart in the reference (a library of two operators).
"""
import time
from collections import OrderedDict, deque
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
    data: torch.Tensor         # A: volume; AT: sinogram; SIRT: sinogram b (in layout z_order)
    angles: torch.Tensor
    SY: int
    SX: int
    n_iter: int = 10
    z_order: int = 0
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
        if self.z_order not in (0, 2):
            raise ValueError(f"z_order must be 0 or 2, got {self.z_order}")

    @property
    def H(self):
        if self.kind == "A":
            return ops.vol_dims(self.data.shape, self.z_order)[2]
        return ops.sino_dims(self.data.shape, self.z_order)[2]

    @property
    def PX(self):
        if self.kind == "A":
            return max(self.SY, self.SX)
        return ops.sino_dims(self.data.shape, self.z_order)[1]

    @property
    def vshape(self):
        return ops.vol_shape(self.SY, self.SX, self.H, self.z_order)

    @property
    def sshape(self):
        return ops.sino_shape(self.n_angles, self.PX, self.H, self.z_order)

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


def _variant(op, r):
    """The variant a single launch of `op` on request r uses (a batch grouped
    by it gives the single-launch results bit for bit)."""
    return int(_lib.load().vp_launch_variant(0 if op == "A" else 1, r.n_angles, r.SY, r.SX, r.H, r.PX))


def _run_chain(r):
    """The plain single-instance operators for one request (on the current
    stream); returns its result."""
    if r.kind == "A":
        return ops.project(r.data, r.angles, r.PX, r.z_order)
    if r.kind == "AT":
        return ops.backproject(r.data, r.angles, r.SY, r.SX, r.z_order)
    solver = recon.SIRT(r.angles, r.SY, r.SX, r.H, r.PX, z_order=r.z_order, device=r.data.device)
    x = torch.zeros(r.vshape, device=r.data.device)
    solver.iterate(x, r.data, r.n_iter)
    return x


class Server:
    """Continuous-batching server on one device.

    mem_budget: bytes of request buffers admitted at a time (default 80 % of
    the free device memory); max_active: request slots; queue_ms / depth:
    estimated GPU time and number of steps queued on the device at most (a
    step's time is estimated from its voxel-angles at `est_rate`; the queue
    always holds two steps when there is work).  Short queues let a new short request
    overtake sooner; long ones hide the host's submission time."""

    def __init__(self, device="cuda", mem_budget=None, max_active=64, grid=None, depth=16,
                 queue_ms=2.0, est_rate=5e12):
        self.device = torch.device(device)
        if mem_budget is None:
            free, _total = torch.cuda.mem_get_info(self.device)
            mem_budget = int(0.8 * free)
        self.mem_budget = int(mem_budget)
        self.max_active = max_active
        self.grid = default_grid(self.device) if grid is None else grid
        self.cus = torch.cuda.get_device_properties(self.device).multi_processor_count
        self.depth = max(1, int(depth))
        self.queue_ms = float(queue_ms)
        self.est_rate = float(est_rate)
        # default-priority streams: two for requests that fill the device, one
        # for the batched small requests
        self.streams = [torch.cuda.Stream(device=self.device) for _ in range(3)]
        self.active = []         # admitted, not retired (arrival order)
        self.in_use = 0
        self.ticks = 0           # steps submitted (batched steps count once)
        self._flight = []        # (event, [requests completing at it])
        self._queued = [[], [], []]  # per stream: (event, estimated ms) of steps still queued or running
        self._plans = OrderedDict()  # cached launch descriptors
        self._ev0 = None

    # ------------------------------------------------------------ admission
    def can_admit(self, req):
        return (len(self.active) < self.max_active
                and self.in_use + req.bytes_needed() <= self.mem_budget)

    def _blocks(self, op, r):
        code = 0 if op == "A" else 1
        return _lib.load().vp_batch_blocks(code, _variant(op, r), r.n_angles, r.SY, r.SX,
                                           r.H, r.PX)

    def fills_device(self, req):
        """Whether the request's own A and AT launches each have at least one
        workgroup per CU (then batching adds nothing to them)."""
        ops_needed = {"A": ("A",), "AT": ("AT",), "SIRT": ("A", "AT")}[req.kind]
        return all(self._blocks(op, req) >= self.cus for op in ops_needed)

    def admit(self, req, now):
        """Allocate the request's buffers (stream-ordered after the data it
        was given) and queue its steps: (op, input, output, post), where post
        is the elementwise work that follows the launch."""
        req.t_admit = now
        self.in_use += req.bytes_needed()
        st = req.state
        small = not self.fills_device(req)
        if small:
            lane = 2
        else:  # the request stream with less remaining work
            left = [sum(r.state["remaining"] for r in self.active if r.state.get("lane") == k)
                    for k in (0, 1)]
            lane = 0 if left[0] <= left[1] else 1
        self.active.append(req)
        stream = self.streams[lane]
        stream.wait_stream(torch.cuda.current_stream(self.device))
        dev = self.device
        steps = deque()
        with torch.cuda.stream(stream):
            if req.kind == "A":
                req.result = torch.empty(req.sshape, device=dev)
                steps.append(("A", req.data, req.result, None))
            elif req.kind == "AT":
                req.result = torch.empty(req.vshape, device=dev)
                steps.append(("AT", req.data, req.result, None))
            else:
                vshape = req.vshape
                sshape = tuple(req.data.shape)
                x = torch.zeros(vshape, device=dev)
                R = torch.empty(sshape, device=dev)
                C = torch.empty(vshape, device=dev)
                ts = torch.empty(sshape, device=dev)
                tv = torch.empty(vshape, device=dev)
                st.update(x=x, R=R, C=C, ts=ts, tv=tv, ones_v=torch.ones(vshape, device=dev),
                          ones_s=torch.ones(sshape, device=dev))
                req.result = x
                b = req.data
                # the same elementwise steps as recon.SIRT
                steps.append(("A", st["ones_v"], R, lambda: _safe_reciprocal_(R)))
                steps.append(("AT", st["ones_s"], C, lambda: _safe_reciprocal_(C)))
                for _ in range(int(req.n_iter)):
                    steps.append(("A", x, ts, lambda: (torch.sub(b, ts, out=ts), ts.mul_(R))))
                    steps.append(("AT", ts, tv, lambda: x.addcmul_(C, tv)))
        st["steps"] = steps
        st["small"] = small
        st["lane"] = lane
        st["remaining"] = req.unit_work() * len(steps)

    # ------------------------------------------------------------ launches
    def _plan(self, op, v, z, jobs):
        """The batched launch of `jobs` [(in, out, angles)] of layout z under
        variant v, built once per set of buffers and reused while that set
        recurs."""
        key = (op, v, z) + tuple((j[0].data_ptr(), j[1].data_ptr(), j[2].data_ptr()) for j in jobs)
        plan = self._plans.get(key)
        if plan is None:
            plan = BatchPlan(op, [j[0] for j in jobs], [j[1] for j in jobs], [j[2] for j in jobs],
                             accumulate=False, variant=v, grid=self.grid, z_order=z)
            self._plans[key] = plan
            while len(self._plans) > 64:
                self._plans.popitem(last=False)
        else:
            self._plans.move_to_end(key)
        return plan

    def _launch_single(self, r, op, inp, out):
        out.zero_()
        if op == "A":
            ops.project(inp, r.angles, r.PX, r.z_order, out=out)
        else:
            ops.backproject(inp, r.angles, r.SY, r.SX, r.z_order, out=out)

    def tick(self, lane=None):
        """Submit one step on one stream (by default the first stream whose
        queue has room and work): the next step of its request with the least
        remaining work, batched with the next steps of the stream's other
        small requests of the same operator.  Nothing waits for the device.
        Returns the requests whose last step it was."""
        if lane is None:
            lanes = [k for k in range(3) if self._ready(k) and self._queue_has_room(k)]
            if not lanes:
                return []
            lane = lanes[0]
        ready = self._ready(lane)
        if not ready:
            return []
        lead = min(ready, key=lambda r: r.state["remaining"])  # ties: arrival order
        op = lead.state["steps"][0][0]
        if lead.state["small"]:
            group = [r for r in ready if r.state["steps"][0][0] == op]
        else:
            group = [lead]
        done = []
        stream = self.streams[lane]
        with torch.cuda.stream(stream):
            steps = [r.state["steps"].popleft() for r in group]
            if len(group) > 1:
                by_variant = {}
                for r, (_, inp, out, _post) in zip(group, steps):
                    by_variant.setdefault((_variant(op, r), r.z_order), []).append((inp, out, r.angles))
                for (v, z), jobs in sorted(by_variant.items()):
                    self._plan(op, v, z, jobs).launch()
            else:
                _, inp, out, _post = steps[0]
                self._launch_single(lead, op, inp, out)
            for r, (_, _inp, _out, post) in zip(group, steps):
                if post is not None:
                    post()
                r.state["remaining"] -= r.unit_work()
                if not r.state["steps"]:
                    done.append(r)
            ev = torch.cuda.Event(enable_timing=True)
            ev.record(stream)
        self.ticks += 1
        self._queued[lane].append((ev, sum(r.unit_work() for r in group) / self.est_rate * 1e3))
        if done:
            self._flight.append((ev, done))
        return done

    def _ready(self, lane):
        return [r for r in self.active if r.state.get("lane") == lane and r.state.get("steps")]

    def _retire(self, r, t_done):
        r.t_done = t_done
        self.in_use -= r.bytes_needed()
        for k in ("R", "C", "ts", "tv", "ones_v", "ones_s", "steps", "lane"):
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
        self._queued = [[(e, ms) for e, ms in q if not e.query()] for q in self._queued]
        return fired

    def _queue_has_room(self, lane):
        # two steps at least, so the next one is already queued when one ends
        q = self._queued[lane]
        return len(q) < 2 or (len(q) < self.depth and sum(ms for _, ms in q) < self.queue_ms)

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
            self._poll()
            if any(self._ready(k) and self._queue_has_room(k) for k in range(3)):
                self.tick()
                continue
            if not self.active:
                if i == len(pending):
                    break
                # idle: wait for the next arrival
                time.sleep(max(0.0, pending[i].arrival - now))
                continue
            # the queue is full: look again shortly (new arrivals, completions)
            time.sleep(5e-5)
        torch.cuda.synchronize(self.device)
        for st in self.streams:
            main.wait_stream(st)
        self.makespan = time.perf_counter() - t0
        return requests


def run_alone(req, device="cuda"):
    """The same request through the same kernels with nothing else in flight."""
    s = Server(device=device)
    clone = Request(req.rid, req.kind, 0.0, req.data, req.angles, req.SY, req.SX, req.n_iter,
                    req.z_order)
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
                   n_angles=(90, 192, 360, 720, 1440), n_iter=10, device="cuda", layouts=(0,)):
    """Poisson request stream: 40% A, 40% AT, 20% SIRT; ellipsoid-phantom or
    N(0,1) inputs generated on the device (seeded); each request's layout
    drawn from `layouts` (the default draws nothing more from the generator,
    so the z_order 0 stream is the same as without layouts)."""
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
        z = int(rng.choice(layouts)) if len(layouts) > 1 else int(layouts[0])
        if z == 2:
            data = data.permute(2, 0, 1).contiguous()
        reqs.append(Request(i, str(kind), t, data, ang, S, S, n_iter, z))
    return reqs


def latency_stats(requests):
    lat = np.array([(r.t_done - r.arrival) * 1e3 for r in requests])
    return {"p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean()), "max_ms": float(lat.max())}


def throughput(requests, seconds):
    return sum(r.work() for r in requests) / seconds / 1e9  # Gvoxel*angle/s
