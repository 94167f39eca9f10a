"""Batched and served projectors in the z_order 2 layout (H slowest).

The batched kernels read the z_order 0 layout; a z_order 2 plan transposes
its inputs into slice-fastest workspaces and its results back (batched
layout transposes), exactly as the single-instance entry points do, so its
results equal the plain z_order 2 operators bit for bit.
"""
import ctypes

import numpy as np
import pytest
import torch

import oracle as vpo
from voxelproj_amd import _lib, batch, ops, serve

pytestmark = pytest.mark.gpu
DEV = "cuda"

# (SY, SX, H, n_angles, PX): one per kernel variant, ragged extents
SHAPES = [(40, 36, 70, 45, 52), (96, 96, 32, 60, 96), (17, 23, 5, 13, 30), (50, 70, 256, 33, 70),
          (64, 64, 129, 90, 64)]


def instances(seed):
    g = torch.Generator(device=DEV).manual_seed(seed)
    rng = np.random.default_rng(seed)
    xs, ys, angs = [], [], []
    for SY, SX, H, A, PX in SHAPES:
        xs.append(torch.randn((H, SY, SX), device=DEV, generator=g))
        ys.append(torch.randn((H, A, PX), device=DEV, generator=g))
        angs.append(torch.tensor(rng.uniform(0, np.pi, A).astype(np.float32), device=DEV))
    return xs, ys, angs


class _T(ctypes.Structure):
    _fields_ = [("inp", ctypes.c_void_p), ("out", ctypes.c_void_p), ("rows", ctypes.c_int64),
                ("cols", ctypes.c_int64), ("accumulate", ctypes.c_int32), ("reserved", ctypes.c_int32)]


def test_transpose_batched_is_exact():
    g = torch.Generator(device=DEV).manual_seed(5)
    dims = [(64, 64), (128, 192), (3, 5), (65, 130), (1, 1000), (1000, 1), (256, 4), (7, 64)]
    ins = [torch.randn(d, device=DEV, generator=g) for d in dims]
    outs = [torch.randn((d[1], d[0]), device=DEV, generator=g) for d in dims]
    olds = [o.clone() for o in outs]
    arr = (_T * len(dims))(*[_T(a.data_ptr(), b.data_ptr(), d[0], d[1], i % 2, 0)
                             for i, (a, b, d) in enumerate(zip(ins, outs, dims))])
    dev = torch.tensor(np.frombuffer(bytes(arr), dtype=np.uint8), device=DEV)
    for grid in (0, 3):
        for o, old in zip(outs, olds):
            o.copy_(old)
        rc = _lib.load().vp_transpose_batched(ctypes.addressof(arr), dev.data_ptr(), len(dims), grid,
                                              torch.cuda.current_stream().cuda_stream)
        _lib.check(rc, "transpose")
        torch.cuda.synchronize()
        for i, (a, o, old) in enumerate(zip(ins, outs, olds)):
            want = a.t() + old if i % 2 else a.t()
            assert torch.equal(o, want), (i, grid)


@pytest.mark.parametrize("op", ["A", "AT"])
def test_z2_batch_equals_plain_z2_and_oracle(op):
    xs, ys, angs = instances(11 if op == "A" else 12)
    if op == "A":
        outs = [torch.full(ops.sino_shape(a.numel(), s[4], s[2], 2), 3.0, device=DEV) for a, s in zip(angs, SHAPES)]
        ins = xs
    else:
        outs = [torch.full((s[2], s[0], s[1]), 3.0, device=DEV) for s in SHAPES]
        ins = ys
    olds = [o.clone() for o in outs]
    gp = batch.GroupedPlan(op, ins, outs, angs, accumulate=True, z_order=2)
    assert len(gp.plans) >= 3  # variants 1, 2 and 4 all present
    gp.launch()
    # overwriting plan (accumulate=False) into fresh outputs
    news = [torch.full_like(o, 9.0) for o in outs]
    batch.GroupedPlan(op, ins, news, angs, accumulate=False, z_order=2).launch()
    fresh = [None] * len(SHAPES)
    for i, (inp, a, s) in enumerate(zip(ins, angs, SHAPES)):
        ref = olds[i].clone()
        if op == "A":
            ops.project(inp, a, s[4], 2, out=ref)
        else:
            ops.backproject(inp, a, s[0], s[1], 2, out=ref)
        assert torch.equal(outs[i], ref), i
        # the z_order 0 batch of the same data, transposed: the same bits
        inp0 = inp.permute(1, 2, 0).contiguous()
        if op == "A":
            out0 = batch.forward_batch([inp0], [a], PX=s[4])[0]
        else:
            out0 = batch.backward_batch([inp0], [a], xs=[torch.empty((s[0], s[1], s[2]), device=DEV)])[0]
        fresh[i] = out0.permute(2, 0, 1)
        if s[0] * s[1] * s[2] * s[3] < 3e7:
            ni = inp.cpu().numpy().astype(np.float64)
            if op == "A":
                want = vpo.forward_f64(ni, a.cpu().numpy(), s[4], 2)
            else:
                want = vpo.backward_f64(ni, a.cpu().numpy(), s[0], s[1], 2)
            got = news[i].cpu().numpy()
            assert float(np.linalg.norm(got - want) / np.linalg.norm(want)) < 1e-6, i
    for i in range(len(SHAPES)):
        assert torch.equal(news[i], fresh[i]), i


def test_z2_graph_pair_replays_exactly():
    xs, _, angs = instances(13)
    ys = [torch.empty(ops.sino_shape(a.numel(), s[4], s[2], 2), device=DEV) for a, s in zip(angs, SHAPES)]
    xb = [torch.empty_like(x) for x in xs]
    pa = batch.GroupedPlan("A", xs, ys, angs, z_order=2)
    pt = batch.GroupedPlan("AT", ys, xb, angs, z_order=2)
    pa.launch()
    pt.launch()
    torch.cuda.synchronize()
    want_y = [y.clone() for y in ys]
    want_x = [x.clone() for x in xb]
    gr = batch.capture_pair(pa, pt)
    for _ in range(3):
        for t in ys + xb:
            t.fill_(-1.0)
        gr.replay()
        torch.cuda.synchronize()
        assert all(torch.equal(a, b) for a, b in zip(ys, want_y))
        assert all(torch.equal(a, b) for a, b in zip(xb, want_x))


def test_served_mixed_layouts_equal_plain():
    reqs = serve.mixed_workload(16, rate=2000.0, seed=7, sizes=(24, 40, 64), heights=(8, 33, 64, 130),
                                n_angles=(20, 45, 90), n_iter=3, layouts=(0, 2))
    assert {r.z_order for r in reqs} == {0, 2}
    server = serve.Server()
    server.run(reqs)
    assert server.in_use == 0 and not server.active
    plain, _ = serve.run_plain(reqs)
    for r, p in zip(reqs, plain):
        want_shape = r.sshape if r.kind == "A" else r.vshape
        assert tuple(r.result.shape) == tuple(want_shape), (r.rid, r.kind, r.z_order)
        assert torch.equal(p, r.result), (r.rid, r.kind, r.z_order)
