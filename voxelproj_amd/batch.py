"""Batched projectors: many volumes of uneven size and angle count per launch.

A ``BatchPlan`` holds the descriptors of a list of instances (host copy for
validation, device copy for the kernel) and launches them as one persistent
grouped kernel through ``vp_forward_batched`` / ``vp_backward_batched``.
The batched kernels read the z_order 0 layout (H fastest); a plan of
z_order 2 instances keeps slice-fastest copies of its inputs and outputs and
wraps the kernel in two batched layout transposes (``vp_transpose_batched``),
as the single-instance entry points do.  Plans only read device memory when
launched, so an A plan followed by an AT plan can be captured into a HIP graph
(``capture_pair``) and replayed.

The kernel variant (H slices per lane, which fixes the tile shapes) is
chosen by ``autotune``: it times every variant on the actual batch and
caches the winner per batch signature, in memory and in a JSON file
(``$VOXELPROJ_AUTOTUNE_CACHE``, default ``~/.cache/voxelproj_amd/autotune.json``).

Every instance's result is independent of the others in the batch and of
the grid size (fixed work order, no atomics): a batch of one and a batch of
many give bitwise identical results for the same variant.

No counterpart in the reference, which projects one volume per call
(voxelproj/wrappers.py:28-122).
"""
import ctypes
import json
import math
import os
import threading

import numpy as np
import torch

from . import _lib, ops, recon

FORWARD_VARIANTS = (1, 2, 4, 416, 418, 420, 32)
BACKWARD_VARIANTS = (1, 2, 4, 32, 64, 128)


class _TItem(ctypes.Structure):
    _fields_ = [
        ("in_", ctypes.c_void_p),
        ("out", ctypes.c_void_p),
        ("rows", ctypes.c_int64),
        ("cols", ctypes.c_int64),
        ("accumulate", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


def _to_device(arr, device):
    """A ctypes descriptor array as (pinned host copy, device copy): the
    upload is stream-ordered from pinned memory, so building a plan while
    earlier work is in flight does not wait for the device."""
    raw = np.frombuffer(bytes(arr), dtype=np.uint8)
    pinned = torch.empty(raw.size, dtype=torch.uint8, pin_memory=True)
    pinned.numpy()[:] = raw
    with torch.cuda.device(device):
        dev = pinned.to(device, non_blocking=True)
    return pinned, dev


class _Item(ctypes.Structure):
    _fields_ = [
        ("in_", ctypes.c_void_p),
        ("out", ctypes.c_void_p),
        ("angles", ctypes.c_void_p),
        ("n_angles", ctypes.c_int32),
        ("SY", ctypes.c_int32),
        ("SX", ctypes.c_int32),
        ("H", ctypes.c_int32),
        ("PX", ctypes.c_int32),
        ("accumulate", ctypes.c_int32),
    ]


def default_grid(device=None):
    """Persistent workgroups: two per compute unit."""
    props = torch.cuda.get_device_properties(device if device is not None else torch.cuda.current_device())
    return 2 * props.multi_processor_count


class BatchPlan:
    """One batched launch of A (op='A') or AT (op='AT') over instances of one
    layout.

    ins / outs: lists of contiguous float32 device tensors; angles: list of
    float32 device tensors.  For A, ins are volumes and outs sinograms; for
    AT the other way round; z_order 0: volumes (SY, SX, H), sinograms
    (A, PX, H); z_order 2: (H, SY, SX) and (H, A, PX).  accumulate=False
    overwrites the outputs.  A z_order 2 plan holds slice-fastest workspaces
    the size of its inputs and outputs (from torch's allocator, kept for the
    plan's lifetime).
    """

    def __init__(self, op, ins, outs, angles, accumulate=False, variant=None, grid=None, z_order=0):
        if op not in ("A", "AT"):
            raise ValueError("op must be 'A' or 'AT'")
        if z_order not in (0, 2):
            raise ValueError(f"z_order must be 0 or 2, got {z_order}")
        self.z_order = z_order
        if not (len(ins) == len(outs) == len(angles)) or len(ins) == 0:
            raise ValueError("ins, outs and angles must be non-empty lists of equal length")
        self.op = op
        self.device = ins[0].device
        self.ins, self.outs = list(ins), list(outs)
        self.angles = [ops.device_angles(a, self.device) for a in angles]
        n = len(ins)
        arr = (_Item * n)()
        t_in, t_out = (_TItem * n)(), (_TItem * n)()
        self.shapes = []
        self._ws = []
        acc = 1 if accumulate else 0
        for i, (a, b, ang) in enumerate(zip(self.ins, self.outs, self.angles)):
            for t, nm in ((a, "input"), (b, "output")):
                if not (t.is_cuda and t.dtype == torch.float32 and t.is_contiguous() and t.dim() == 3):
                    raise ValueError(f"instance {i}: {nm} must be a contiguous 3-D float32 device tensor")
                if t.device != self.device:
                    raise ValueError(f"instance {i}: all tensors must be on {self.device}")
            if op == "A":
                SY, SX, H = ops.vol_dims(a.shape, z_order)
                A_, PX, H2 = ops.sino_dims(b.shape, z_order)
            else:
                A_, PX, H = ops.sino_dims(a.shape, z_order)
                SY, SX, H2 = ops.vol_dims(b.shape, z_order)
            if H2 != H or A_ != ang.numel():
                raise ValueError(f"instance {i}: inconsistent shapes {tuple(a.shape)} / {tuple(b.shape)} "
                                 f"with {ang.numel()} angles")
            if z_order == 0:
                arr[i] = _Item(a.data_ptr(), b.data_ptr(), ang.data_ptr(), A_, SY, SX, H, PX, acc)
            else:
                # (H, rest) -> (rest, H) into the workspace, the kernel on the
                # workspaces, then (rest, H) -> (H, rest) added into the output
                wi = torch.empty((a.numel() // H, H), device=self.device)
                wo = torch.empty((b.numel() // H, H), device=self.device)
                self._ws += [wi, wo]
                arr[i] = _Item(wi.data_ptr(), wo.data_ptr(), ang.data_ptr(), A_, SY, SX, H, PX, 0)
                t_in[i] = _TItem(a.data_ptr(), wi.data_ptr(), H, a.numel() // H, 0, 0)
                t_out[i] = _TItem(wo.data_ptr(), b.data_ptr(), b.numel() // H, H, acc, 0)
            self.shapes.append((int(SY), int(SX), int(H), int(A_), int(PX)))
        self._host = arr
        self._pinned, self._dev = _to_device(arr, self.device)
        if z_order == 2:
            self._t_host = (t_in, t_out)
            self._t_dev = (_to_device(t_in, self.device), _to_device(t_out, self.device))
        self.n = n
        self.grid = default_grid(self.device) if grid is None else int(grid)
        self.variant = variant if variant is not None else _default_variant(op, self.shapes)

    @property
    def work(self):
        """voxel-angles of the batch (SY*SX*H*A summed over instances)."""
        return sum(SY * SX * H * A for SY, SX, H, A, _ in self.shapes)

    def signature(self):
        return f"{self.op}:z{self.z_order}:" + ";".join(",".join(map(str, s)) for s in sorted(self.shapes))

    def _transpose(self, k, stream):
        lib = _lib.load()
        rc = lib.vp_transpose_batched(ctypes.addressof(self._t_host[k]), self._t_dev[k][1].data_ptr(),
                                      self.n, 0, stream)
        _lib.check(rc, "batched layout transpose")

    def launch(self, variant=None):
        lib = _lib.load()
        v = self.variant if variant is None else variant
        fn = lib.vp_forward_batched if self.op == "A" else lib.vp_backward_batched
        with torch.cuda.device(self.device):
            stream = torch.cuda.current_stream(self.device).cuda_stream
            if self.z_order == 2:
                self._transpose(0, stream)
            rc = fn(int(v), ctypes.addressof(self._host), self._dev.data_ptr(), self.n,
                    self.grid, stream)
            _lib.check(rc, "batched " + self.op)
            if self.z_order == 2:
                self._transpose(1, stream)
        return self.outs


def _default_variant(op, shapes):
    """The default of the median H; for a plan of one instance, the variant
    its single launch uses (the same bits)."""
    lib = _lib.load()
    if len(shapes) == 1:
        SY, SX, H, A_, PX = shapes[0]
        return int(lib.vp_launch_variant(0 if op == "A" else 1, A_, SY, SX, H, PX))
    hs = sorted(s[2] for s in shapes)
    return int(lib.vp_default_variant(0 if op == "A" else 1, hs[len(hs) // 2]))


# --------------------------------------------------------------- autotuner
_cache = {}
_cache_lock = threading.Lock()


def _cache_path():
    return os.environ.get("VOXELPROJ_AUTOTUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "voxelproj_amd",
                                       "autotune.json"))


def _load_cache():
    p = _cache_path()
    try:
        with open(p) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store_cache(entries):
    p = _cache_path()
    try:
        os.makedirs(os.path.dirname(p), exist_ok=True)
        tmp = p + ".tmp"
        with open(tmp, "w") as f:
            json.dump(entries, f, indent=1, sort_keys=True)
        os.replace(tmp, p)
    except OSError:
        pass  # a read-only home only costs re-tuning


def autotune(plan, reps=3, persist=True):
    """Time every kernel variant on `plan` (outputs are overwritten or
    accumulated into as the plan says, so use scratch outputs), pick the
    fastest, store it in the plan and the cache, and return {variant: ms}."""
    key = f"{torch.cuda.get_device_name(plan.device)}|{plan.signature()}|grid={plan.grid}"
    with _cache_lock:
        if not _cache:
            _cache.update(_load_cache())
        if key in _cache:
            plan.variant = int(_cache[key]["best"])
            return {int(k): v for k, v in _cache[key]["ms"].items()}
    variants = FORWARD_VARIANTS if plan.op == "A" else BACKWARD_VARIANTS
    times = {}
    stream = torch.cuda.current_stream(plan.device)
    for v in variants:
        plan.launch(v)  # warm-up (code object load, caches)
        start = torch.cuda.Event(enable_timing=True)
        end = torch.cuda.Event(enable_timing=True)
        start.record(stream)
        for _ in range(reps):
            plan.launch(v)
        end.record(stream)
        end.synchronize()
        times[v] = start.elapsed_time(end) / reps
    best = min(times, key=times.get)
    plan.variant = best
    with _cache_lock:
        _cache[key] = {"best": best, "ms": {str(k): t for k, t in times.items()}}
        if persist:
            _store_cache(dict(_cache))
    return times


class GroupedPlan:
    """A batch split into groups by kernel variant, one persistent launch per
    group.  One variant for a whole uneven batch leaves lanes idle on the
    instances whose H it does not suit (4 slices per lane on H = 32 leaves 7
    lanes in 8 idle); grouping gives every instance the variant its H calls
    for -- the one the single-instance operator uses, so each result equals
    the plain operator's bit for bit -- while each group still runs as one
    persistent launch.  ``variant_of(H)`` overrides the grouping rule.
    Forward instances above 128 slices take the strip width their single
    launch would (``vp_launch_variant``: 4, 416, 418 or 420, by how well the
    instance's own grid fills the device), for the same bits.
    Forward instances with H <= 32 form a group of their own (variant 1 by
    default, like the single-instance operator), so that ``autotune_groups``
    can give them variant 32 (two angles per sweep wave: no idle lanes at
    H <= 32) without changing the variant of the H = 33..64 instances.
    """

    def __init__(self, op, ins, outs, angles, accumulate=False, grid=None, variant_of=None,
                 z_order=0):
        if not (len(ins) == len(outs) == len(angles)) or len(ins) == 0:
            raise ValueError("ins, outs and angles must be non-empty lists of equal length")
        if z_order not in (0, 2):
            raise ValueError(f"z_order must be 0 or 2, got {z_order}")
        lib = _lib.load()
        code = 0 if op == "A" else 1
        groups = {}
        for i, t in enumerate(ins):
            # H: the last extent of volumes and sinograms in z_order 0, the first in 2
            H = int(t.shape[2] if z_order == 0 else t.shape[0])
            v = int(variant_of(H)) if variant_of is not None else int(lib.vp_default_variant(code, H))
            if variant_of is None and op == "A" and v == 4:
                SY, SX, _ = ops.vol_dims(t.shape, z_order)
                A_, PX, _ = ops.sino_dims(outs[i].shape, z_order)
                v = int(lib.vp_launch_variant(0, int(A_), int(SY), int(SX), H, int(PX)))
            narrow = variant_of is None and op == "A" and H <= 32
            groups.setdefault((v, narrow), []).append(i)
        self.op = op
        self.device = ins[0].device
        self.plans = [BatchPlan(op, [ins[i] for i in idx], [outs[i] for i in idx],
                                [angles[i] for i in idx], accumulate, v, grid, z_order)
                      for (v, _), idx in sorted(groups.items())]
        self.outs = list(outs)
        self._streams = None

    @property
    def work(self):
        return sum(p.work for p in self.plans)

    @property
    def variants(self):
        """{variant: number of instances} over the groups."""
        out = {}
        for p in self.plans:
            out[p.variant] = out.get(p.variant, 0) + p.n
        return out

    def launch(self, concurrent=False):
        """One launch per group on the current stream; with concurrent=True
        the groups run on side streams (default priority) forked from and
        joined back into the current stream, so the workgroups of one group
        can fill the CUs left idle by another's tail.  The groups write
        disjoint outputs, so the results do not change."""
        if not concurrent or len(self.plans) < 2:
            for p in self.plans:
                p.launch()
            return self.outs
        cur = torch.cuda.current_stream(self.device)
        if self._streams is None:
            self._streams = [torch.cuda.Stream(device=self.device) for _ in self.plans]
        for p, st in zip(self.plans, self._streams):
            st.wait_stream(cur)
            with torch.cuda.stream(st):
                p.launch()
        for st in self._streams:
            cur.wait_stream(st)
        return self.outs


def autotune_grouped(gplan, reps=3, persist=True):
    """Autotune every group of a GroupedPlan; returns [(group size, {variant: ms})]."""
    return [(p.n, autotune(p, reps, persist)) for p in gplan.plans]


# --------------------------------------------------------------- graphs
def capture_pair(plan_a, plan_at, warmup=1, concurrent=False):
    """Capture `plan_a.launch()` followed by `plan_at.launch()` into a graph.

    Returns a torch.cuda.CUDAGraph; replay() re-runs both launches on the
    same buffers.  Use accumulate=False plans (no fill kernels needed).  With
    concurrent=True, GroupedPlans capture their groups as parallel branches
    (AT still after all of A)."""
    def run():
        if concurrent and isinstance(plan_a, GroupedPlan):
            plan_a.launch(concurrent=True)
        else:
            plan_a.launch()
        if concurrent and isinstance(plan_at, GroupedPlan):
            plan_at.launch(concurrent=True)
        else:
            plan_at.launch()

    s = torch.cuda.Stream(device=plan_a.device)
    s.wait_stream(torch.cuda.current_stream(plan_a.device))
    with torch.cuda.stream(s):
        for _ in range(warmup):
            run()
    torch.cuda.current_stream(plan_a.device).wait_stream(s)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        run()
    return g


# --------------------------------------------------------------- workload
def uneven_workload(n=48, seed=0, device="cuda"):
    """The batched benchmark set: n volumes with SX = SY from
    {96, 128, 192, 256, 384}, H from {32, 64, 128, 256}, n_angles from
    {60, 180, 360, 720, 1024}, angles uniform in [0, pi) with +-0.5 degree
    seeded jitter; half N(0,1) volumes, half ellipsoid phantoms.

    Returns a list of (x, angles) device tensors (x in z_order 0)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        S = int(rng.choice([96, 128, 192, 256, 384]))
        H = int(rng.choice([32, 64, 128, 256]))
        A = int(rng.choice([60, 180, 360, 720, 1024]))
        ang = np.linspace(0, np.pi, A, endpoint=False) + rng.uniform(-0.5, 0.5, A) * math.pi / 180
        ang_t = torch.tensor(ang.astype(np.float32), device=device)
        if i % 2 == 0:
            g = torch.Generator(device=device).manual_seed(seed * 1000 + i)
            x = torch.randn((S, S, H), device=device, generator=g)
        else:
            x = recon.ellipsoid_phantom(S, S, H, seed=seed * 1000 + i, device=device)
        out.append((x, ang_t))
    return out


def forward_batch(xs, angles, ys=None, PX=None, accumulate=False, variant=None, grid=None,
                  z_order=0):
    """A for a list of volumes in one launch; returns the list of sinograms."""
    if ys is None:
        ys = []
        for x, a in zip(xs, angles):
            SY, SX, H = ops.vol_dims(x.shape, z_order)
            px = max(SY, SX) if PX is None else PX
            ys.append(torch.empty(ops.sino_shape(ops.device_angles(a, x.device).numel(), px, H, z_order),
                                  device=x.device, dtype=torch.float32))
        accumulate = False
    BatchPlan("A", xs, ys, angles, accumulate, variant, grid, z_order).launch()
    return ys


def backward_batch(ys, angles, xs=None, accumulate=False, variant=None, grid=None, z_order=0):
    """AT for a list of sinograms in one launch; returns the list of volumes."""
    if xs is None:
        xs = []
        for y, a in zip(ys, angles):
            A_, PX, H = ops.sino_dims(y.shape, z_order)
            xs.append(torch.empty(ops.vol_shape(PX, PX, H, z_order), device=y.device, dtype=torch.float32))
        accumulate = False
    BatchPlan("AT", ys, xs, angles, accumulate, variant, grid, z_order).launch()
    return xs
