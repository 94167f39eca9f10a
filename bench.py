"""Benchmark of the exact-trapezoid projector pair on one MI355X.

A step is one forward projection A and one backprojection AT of the
plain-operator headline workload: float32 volume x ~ N(0,1), shape
(SY, SX, H) = (512, 512, 512) in z_order 0, 192 angles linspace(0, pi),
PX = 512 (reference default max(SY, SX)), preallocated outputs, inputs
resident in HBM.  value = voxel-angles processed per second by the step
(2 * 512^3 * 192 per step) in Gvoxel*angle/s.

Also reported:
  roofline      for the dominant kernel, from HIP-event timings on the
                stream the kernels run on; algorithmic flops = 2*H*nnz(W)
                (one FMA per nonzero weight per slice), nnz counted exactly
  per_op        A and AT rates, and both ops in z_order 2
  cpu_baseline  the OpenMP FP32 transcription of the reference algorithm
                (oracle/, "port") timed on this host on a bounded sample
  traffic       HBM bytes per launch when a PMC summary is supplied
  batched       48 uneven volumes: plain one-at-a-time A+AT vs the autotuned
                persistent batched A+AT replayed from a HIP graph
  serving       Poisson stream of A / AT / SIRT-10 requests: continuous
                batching (p50/p99 latency, throughput) vs one at a time

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np
import torch

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3    # MI355X published FP32 vector peak (packed)
PEAK_HBM_GBS = 8000.0       # MI355X HBM3E


def count_nnz(angles, SY, SX, PX):
    """Exact number of nonzero weights: pixels p in [0, PX) with |p - u| < b + 1/2."""
    iy, ix = np.meshgrid(np.arange(SY, dtype=np.float64), np.arange(SX, dtype=np.float64),
                         indexing="ij")
    X = (ix - 0.5 * (SX - 1)).ravel()
    Y = (iy - 0.5 * (SY - 1)).ravel()
    pc = 0.5 * (PX - 1)
    total = 0
    for phi in np.asarray(angles, np.float32).astype(np.float64):
        c, s = np.cos(phi), np.sin(phi)
        half = 0.5 * (abs(c) + abs(s)) + 0.5
        u = X * c + Y * s + pc
        lo = np.maximum(np.floor(u - half) + 1, 0)
        hi = np.minimum(np.ceil(u + half) - 1, PX - 1)
        total += int(np.maximum(hi - lo + 1, 0).sum())
    return total


def time_op(fn, args, n, stream, warmup=3):
    """Per-call launch durations (ms) of n back-to-back calls, each bracketed
    by HIP events on the kernels' stream, after `warmup` untimed calls.
    Returns (median, p10, p90)."""
    for _ in range(warmup):
        fn(*args)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(n + 1)]
    ev[0].record(stream)
    for i in range(n):
        fn(*args)
        ev[i + 1].record(stream)
    ev[-1].synchronize()
    t = np.array([ev[i].elapsed_time(ev[i + 1]) for i in range(n)])
    return float(np.median(t)), float(np.percentile(t, 10)), float(np.percentile(t, 90))


def cpu_model():
    """The host CPU's model name (from /proc/cpuinfo), or the platform string."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or platform.machine()


def _time_pair(fwd, bwd, budget_s):
    """Best (total, A, AT) seconds over repetitions of fwd(); bwd() until
    about budget_s seconds are spent (at least one repetition)."""
    best, spent, reps = None, 0.0, 0
    while spent < budget_s or reps < 1:
        t0 = time.perf_counter()
        fwd()
        t1 = time.perf_counter()
        bwd()
        t2 = time.perf_counter()
        spent += t2 - t0
        reps += 1
        if best is None or t2 - t0 < best[0]:
            best = (t2 - t0, t1 - t0, t2 - t1)
    return best, spent, reps


def cpu_baseline(N, A, H_sample, budget_s=10.0):
    """The reference algorithm (FP32 transcription, OpenMP) and the exact FP64
    oracle on H_sample slices of the workload, each repeated for about
    budget_s seconds of CPU time; rates from the best repetition.  The FP32
    transcription is the reported value."""
    import oracle as vpo

    g = torch.Generator().manual_seed(5)
    ang = np.linspace(0, np.pi, A).astype(np.float32)
    x = torch.randn((N, N, H_sample), generator=g).numpy()
    y = torch.randn((A, N, H_sample), generator=g).numpy()
    vpo.ref_forward_f32(x[:, :, :1], ang, N, 0)  # warm up threads / pages
    b32, s32, r32 = _time_pair(lambda: vpo.ref_forward_f32(x, ang, N, 0),
                               lambda: vpo.ref_backward_f32(y, ang, N, N, 0), budget_s)
    b64, s64, r64 = _time_pair(lambda: vpo.forward_f64(x, ang, N, 0),
                               lambda: vpo.backward_f64(y, ang, N, N, 0), budget_s)
    work = 2.0 * N * N * H_sample * A
    return {
        "value": round(work / b32[0] / 1e9, 3),
        "unit": "Gvoxel*angle/s",
        "cores": vpo.max_threads(),
        "kind": "port",
        "cpu_model": cpu_model(),
        "fp64_oracle_value": round(work / b64[0] / 1e9, 3),
        "sample": (f"A then AT on {N}x{N}x{H_sample} slices x {A} angles (z_order 0), OpenMP on "
                   f"{vpo.max_threads()} threads: FP32 transcription of the reference kernels "
                   f"(oracle/), {r32} repetitions in {s32:.1f} s, best A {b32[1]:.2f} s, AT "
                   f"{b32[2]:.2f} s; exact FP64 oracle, {r64} repetitions in {s64:.1f} s, best A "
                   f"{b64[1]:.2f} s, AT {b64[2]:.2f} s"),
    }


# The reference's published single-op times (README.md:58-59, RTX 3060,
# z_order 0, 192 angles, ms) for the cubic sizes it reports.
REFERENCE_MS = {128: (8.5, 7.7), 256: (59.8, 50.3), 512: (516.0, 397.0)}


def bench_configs(dev, stream, n_t=20):
    """A and AT (z_order 0) alone at the other workload sizes: 128^3 and 256^3
    x 192 angles (linspace(0, pi)) and 1024^2 x 512 slices x 1800 angles
    (linspace(0, pi) without the endpoint), PX = max(SY, SX); median of n_t
    calls (HIP events)."""
    from voxelproj_amd import _CU
    res = {}
    for SY, SX, H, A, endpoint in ((128, 128, 128, 192, True), (256, 256, 256, 192, True),
                                   (1024, 1024, 512, 1800, False)):
        PX = max(SY, SX)
        ang = torch.tensor(np.linspace(0, np.pi, A, endpoint=endpoint).astype(np.float32), device=dev)
        g = torch.Generator(device=dev).manual_seed(1)
        x = torch.randn((SY, SX, H), device=dev, generator=g)
        yb = torch.randn((A, PX, H), device=dev, generator=g)
        y = torch.zeros((A, PX, H), device=dev)
        xb = torch.zeros((SY, SX, H), device=dev)
        n = n_t if SY * SX * H * A < 1e11 else 5
        ms_a = time_op(_CU.forward_z0, (x, y, ang), n, stream)[0]
        ms_at = time_op(_CU.backward_z0, (yb, xb, ang), n, stream)[0]
        va = float(SY) * SX * H * A
        key = f"{SY}x{SX}x{H}x{A}"
        res[key] = {"A_ms": round(ms_a, 4), "AT_ms": round(ms_at, 4),
                    "A_gvox_ang_s": round(va / ms_a / 1e6, 1),
                    "AT_gvox_ang_s": round(va / ms_at / 1e6, 1), "timed_calls": n}
        if SY == SX == H and A == 192 and SY in REFERENCE_MS:
            ra, rat = REFERENCE_MS[SY]
            res[key]["reference_rtx3060_ms"] = [ra, rat]
        del x, yb, y, xb
        torch.cuda.synchronize()
    return res


def bench_batched(dev, n=48, seed=0, reps=3):
    """Batched form (uneven volumes): plain one-instance-at-a-time A then AT vs
    the autotuned persistent batched A then AT replayed from a HIP graph."""
    from voxelproj_amd import batch, ops

    inst = batch.uneven_workload(n, seed=seed, device=dev)
    xs = [x for x, _ in inst]
    angs = [a for _, a in inst]
    ys = [torch.empty((a.numel(), max(x.shape[0], x.shape[1]), x.shape[2]), device=dev)
          for x, a in inst]
    xb = [torch.empty_like(x) for x in xs]
    work = 2.0 * sum(x.shape[0] * x.shape[1] * x.shape[2] * a.numel() for x, a in inst)
    stream = torch.cuda.current_stream(dev)

    def plain():
        for x, a, y, b in zip(xs, angs, ys, xb):
            y.zero_()
            ops.project(x, a, y.shape[1], out=y)
            b.zero_()
            ops.backproject(y, a, x.shape[0], x.shape[1], out=b)

    ms_plain = time_op(plain, (), reps, stream, warmup=1)[0]
    # one persistent launch per kernel-variant group, each group autotuned
    pa = batch.GroupedPlan("A", xs, ys, angs)
    pb = batch.GroupedPlan("AT", ys, xb, angs)
    tune_a = batch.autotune_grouped(pa, persist=False)
    tune_b = batch.autotune_grouped(pb, persist=False)
    graph = batch.capture_pair(pa, pb)
    ms_batch = time_op(graph.replay, (), reps, stream, warmup=1)[0]
    # the same with the variant groups as parallel graph branches
    graph_c = batch.capture_pair(pa, pb, concurrent=True)
    ms_conc = time_op(graph_c.replay, (), reps, stream, warmup=1)[0]
    return {
        "workload": f"{n} volumes, SX=SY in {{96..384}}, H in {{32..256}}, "
                    f"n_angles in {{60..1024}}, A then AT (z_order 0)",
        "plain_ms": round(ms_plain, 3),
        "batched_graph_ms": round(ms_batch, 3),
        "batched_graph_concurrent_groups_ms": round(ms_conc, 3),
        "plain_gvox_ang_s": round(work / ms_plain / 1e6, 1),
        "batched_gvox_ang_s": round(work / ms_batch / 1e6, 1),
        "speedup": round(ms_plain / ms_batch, 3),
        "groups": {"A": {str(v): n for v, n in pa.variants.items()},
                   "AT": {str(v): n for v, n in pb.variants.items()}},
        "autotune_ms": {"A": [{str(k): round(v, 3) for k, v in t.items()} for _, t in tune_a],
                        "AT": [{str(k): round(v, 3) for k, v in t.items()} for _, t in tune_b]},
    }


def bench_serving(dev, n=40, rate=200.0, seed=0):
    """Serving form: Poisson stream (rate req/s) of 40% A / 40% AT / 20% SIRT-10
    requests; continuous batching vs one request at a time (throughput: back
    to back from time 0; latency: the same arrivals served first-in
    first-out without batching)."""
    from voxelproj_amd import serve

    reqs = serve.mixed_workload(n, rate=rate, seed=seed, device=dev)
    # warm-up: every kernel configuration the mix can use (small batched and
    # single launches at 1, 2 and 4 slices per lane)
    warm = serve.mixed_workload(12, rate=rate, seed=seed + 1, sizes=(64, 256), heights=(64, 128, 256),
                                n_angles=(90,), n_iter=1, device=dev)
    serve.Server(device=dev).run(warm)
    server = serve.Server(device=dev)
    server.run(reqs)
    st = serve.latency_stats(reqs)
    _, t_plain = serve.run_plain(reqs)
    fifo = serve.run_plain_fifo(reqs)
    served = serve.throughput(reqs, server.makespan)
    plain = serve.throughput(reqs, t_plain)
    return {
        "workload": f"{n} requests, Poisson {rate:g} req/s, 40% A / 40% AT / 20% SIRT-10, "
                    f"SX=SY in {{128,256,512}}, H in {{64..512}}, n_angles in {{90..1440}}",
        "p50_ms": round(st["p50_ms"], 2),
        "p99_ms": round(st["p99_ms"], 2),
        "makespan_s": round(server.makespan, 3),
        "ticks": server.ticks,
        "served_gvox_ang_s": round(served, 1),
        "plain_gvox_ang_s": round(plain, 1),
        "speedup": round(served / plain, 3),
        # the same arrival stream served one request at a time (no batching)
        "fifo_plain_p50_ms": round(fifo["p50_ms"], 2),
        "fifo_plain_p99_ms": round(fifo["p99_ms"], 2),
        "fifo_plain_makespan_s": round(fifo["makespan_s"], 3),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--angles", type=int, default=192)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-slices", type=int, default=256)
    ap.add_argument("--profile", action="store_true",
                    help="short run for rocprofv3 counters: one 512 MiB copy (reference "
                         "traffic), then one A and one AT launch; prints no JSON")
    ap.add_argument("--profile-z2", action="store_true",
                    help="short run for rocprofv3: one A and one AT in each layout (z_order 0, "
                         "then 2), twice; prints no JSON")
    ap.add_argument("--no-z2", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-configs", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the projector path is not sharded across GPUs "
              "(it would shard along H, where slices are independent and need no exchange). "
              "Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import voxelproj_amd as vp
    from voxelproj_amd import _CU

    vp.load()
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    N, A = args.size, args.angles
    PX = N
    ang_np = np.linspace(0, np.pi, A).astype(np.float32)
    ang = torch.tensor(ang_np, device=dev)
    g = torch.Generator(device=dev).manual_seed(0)
    x = torch.randn((N, N, N), device=dev, generator=g)
    yb = torch.randn((A, PX, N), device=dev, generator=g)
    y = torch.zeros((A, PX, N), device=dev)
    xb = torch.zeros((N, N, N), device=dev)
    stream = torch.cuda.current_stream(dev)

    def step():
        _CU.forward_z0(x, y, ang)
        _CU.backward_z0(yb, xb, ang)

    if args.profile_z2:
        x2 = x.permute(2, 0, 1).contiguous()
        yb2 = yb.permute(2, 0, 1).contiguous()
        y2 = torch.zeros((N, A, PX), device=dev)
        xb2 = torch.zeros((N, N, N), device=dev)
        for _ in range(2):
            step()
            _CU.forward_z2(x2, y2, ang)
            _CU.backward_z2(yb2, xb2, ang)
        torch.cuda.synchronize()
        print("profile run done: A z0, AT z0, A z2, AT z2, twice", flush=True)
        return

    if args.profile:
        step()
        torch.cuda.synchronize()
        src = torch.empty(128 * 1024 * 1024, device=dev)  # 512 MiB
        src.fill_(1.0)
        dst = torch.empty_like(src)
        dst.copy_(src)                                     # reads 512 MiB, writes 512 MiB
        step()
        torch.cuda.synchronize()
        print("profile run done: copy 512 MiB, A z0, AT z0", flush=True)
        return

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_step = (t1 - t0) * 1e3 / args.steps
    vox_ang = float(N) * N * N * A
    value = 2.0 * vox_ang / (ms_step * 1e-3) / 1e9

    # Per-kernel launch durations (HIP events on the kernels' stream): median
    # of 20 calls after 3 warm-ups, with p10 / p90.
    n_t = 20
    ms_a, a10, a90 = time_op(_CU.forward_z0, (x, y, ang), n_t, stream)
    ms_at, t10, t90 = time_op(_CU.backward_z0, (yb, xb, ang), n_t, stream)
    per_op = {"A_z0_ms": round(ms_a, 4), "A_z0_ms_p10_p90": [round(a10, 4), round(a90, 4)],
              "AT_z0_ms": round(ms_at, 4), "AT_z0_ms_p10_p90": [round(t10, 4), round(t90, 4)],
              "A_z0_gvox_ang_s": round(vox_ang / ms_a / 1e6, 1),
              "AT_z0_gvox_ang_s": round(vox_ang / ms_at / 1e6, 1), "timed_calls": n_t}
    if not args.no_z2:
        del y, xb
        x2 = x.permute(2, 0, 1).contiguous()
        del x
        y2 = torch.zeros((N, A, PX), device=dev)
        yb2 = yb.permute(2, 0, 1).contiguous()
        del yb
        xb2 = torch.zeros((N, N, N), device=dev)
        ms_a2 = time_op(_CU.forward_z2, (x2, y2, ang), n_t, stream)[0]
        ms_at2 = time_op(_CU.backward_z2, (yb2, xb2, ang), n_t, stream)[0]
        per_op.update({"A_z2_ms": round(ms_a2, 4), "AT_z2_ms": round(ms_at2, 4),
                       "A_z2_gvox_ang_s": round(vox_ang / ms_a2 / 1e6, 1),
                       "AT_z2_gvox_ang_s": round(vox_ang / ms_at2 / 1e6, 1)})

    nnz = count_nnz(ang_np, N, N, PX)
    flops = 2.0 * N * nnz
    dom_name, dom_ms = ("forward_z0 (A)", ms_a) if ms_a >= ms_at else ("backproject z0 (AT)", ms_at)
    achieved = flops / (dom_ms * 1e-3) / 1e12
    bytes_a = 4.0 * (N ** 3 + 2.0 * A * PX * N)
    bytes_at = 4.0 * (A * PX * N + 2.0 * N ** 3)
    dom_bytes = bytes_a if ms_a >= ms_at else bytes_at
    roofline = {
        "bound": "valu",
        "kernel": dom_name,
        "achieved": round(achieved, 3),
        "peak": PEAK_FP32_TFLOPS,
        "unit": "TFLOP/s",
        "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
        "traffic": None,
        "flops_per_launch": flops,
        "nnz_per_voxel_angle": round(nnz / (N * N * A), 5),
        "hbm_algorithmic_GBs": round(dom_bytes / (dom_ms * 1e-3) / 1e9, 1),
        "hbm_frac": round(dom_bytes / (dom_ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
    }
    # HBM bytes per launch measured with rocprofv3 PMC runs (FETCH_SIZE, WRITE_SIZE,
    # scaled by a copy of known size); recorded for this workload only.
    traffic_file = os.path.join(ROOT, "profiles", "traffic_512cube_a192.txt")
    if N == 512 and A == 192 and os.path.exists(traffic_file):
        tr = {}
        with open(traffic_file) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 2 and not line.startswith("#"):
                    tr[parts[0]] = float(parts[1])
        roofline["traffic"] = tr.get("forward_z0" if ms_a >= ms_at else "backproject_z0")

    out = {
        "metric": "A & AT Gvoxel*angle/s at 512^3 x 192 angles (plain operator)",
        "value": round(value, 2),
        "unit": "Gvoxel*angle/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded N(0,1) volume and sinogram)",
        "config": {"workload": f"A then AT, z_order 0, {N}^3 volume, {A} angles linspace(0,pi), "
                               f"PX={PX}", "size": N, "angles": A, "PX": PX},
        "roofline": roofline,
        "per_op": per_op,
    }
    if not args.no_configs:
        out["other_sizes"] = bench_configs(dev, stream)
    if not args.no_batched:
        out["batched"] = bench_batched(dev)
    if not args.no_serving:
        out["serving"] = bench_serving(dev)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(N, A, args.cpu_slices)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
