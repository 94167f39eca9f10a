"""Batched persistent launches, autotuner and graph capture on the GPU."""
import json

import numpy as np
import pytest
import torch

import oracle as vpo
from voxelproj_amd import batch, ops

pytestmark = pytest.mark.gpu
DEV = "cuda"

# (SY, SX, H, n_angles, PX): uneven sizes, ragged H, tiny and truncated cases
SHAPES = [(40, 36, 70, 45, 52), (96, 96, 32, 60, 96), (17, 23, 5, 13, 30), (64, 64, 128, 90, 64),
          (128, 128, 64, 180, 128), (33, 20, 9, 600, 25), (1, 1, 3, 4, 2), (50, 70, 256, 33, 70)]


def make_instances(seed=0):
    g = torch.Generator(device=DEV).manual_seed(seed)
    rng = np.random.default_rng(seed)
    xs, ys, angs = [], [], []
    for SY, SX, H, A, PX in SHAPES:
        xs.append(torch.randn((SY, SX, H), device=DEV, generator=g))
        ys.append(torch.randn((A, PX, H), device=DEV, generator=g))
        angs.append(torch.tensor(rng.uniform(0, np.pi, A).astype(np.float32), device=DEV))
    return xs, ys, angs


def rel(a, b):
    return float((a.double() - b.double()).norm() / b.double().norm())


@pytest.mark.parametrize("variant", batch.FORWARD_VARIANTS)
def test_forward_batch_matches_single_and_oracle(variant):
    xs, _, angs = make_instances(1)
    outs = [torch.full((a.numel(), s[4], s[2]), 7.0, device=DEV) for a, s in zip(angs, SHAPES)]
    batch.BatchPlan("A", xs, outs, angs, accumulate=False, variant=variant).launch()
    for i, (x, a, s) in enumerate(zip(xs, angs, SHAPES)):
        ref = ops.project(x, a, s[4])
        assert rel(outs[i], ref) < 1e-6, i
        # batch of one with the same variant: bitwise identical
        solo = torch.empty_like(outs[i])
        batch.BatchPlan("A", [x], [solo], [a], accumulate=False, variant=variant).launch()
        assert torch.equal(solo, outs[i]), i
        if s[0] * s[1] * s[2] * s[3] < 2e6:
            yo = vpo.forward_f64(x.cpu().numpy(), a.cpu().numpy(), s[4], 0)
            assert float(np.linalg.norm(outs[i].cpu().numpy() - yo) / np.linalg.norm(yo)) < 1e-6


@pytest.mark.parametrize("variant", batch.BACKWARD_VARIANTS)
def test_backward_batch_matches_single_and_oracle(variant):
    xs, ys, angs = make_instances(2)
    outs = [torch.zeros_like(x) for x in xs]
    batch.BatchPlan("AT", ys, outs, angs, accumulate=True, variant=variant).launch()
    for i, (y, a, s) in enumerate(zip(ys, angs, SHAPES)):
        ref = ops.backproject(y, a, s[0], s[1])
        assert rel(outs[i], ref) < 1e-6, i
        solo = torch.zeros_like(outs[i])
        batch.BatchPlan("AT", [y], [solo], [a], accumulate=True, variant=variant).launch()
        assert torch.equal(solo, outs[i]), i
        if s[0] * s[1] * s[2] * s[3] < 2e6:
            xo = vpo.backward_f64(y.cpu().numpy(), a.cpu().numpy(), s[0], s[1], 0)
            assert float(np.linalg.norm(outs[i].cpu().numpy() - xo) / np.linalg.norm(xo)) < 1e-6


def test_grid_size_does_not_change_results():
    xs, ys, angs = make_instances(3)
    res = []
    for grid in (1, 7, None):
        outs = [torch.empty((a.numel(), s[4], s[2]), device=DEV) for a, s in zip(angs, SHAPES)]
        batch.BatchPlan("A", xs, outs, angs, grid=grid).launch()
        res.append(outs)
    for r in res[1:]:
        assert all(torch.equal(a, b) for a, b in zip(res[0], r))


def test_autotune_picks_fastest_and_caches(tmp_path, monkeypatch):
    monkeypatch.setenv("VOXELPROJ_AUTOTUNE_CACHE", str(tmp_path / "tune.json"))
    batch._cache.clear()
    xs, ys, angs = make_instances(4)
    outs = [torch.empty((a.numel(), s[4], s[2]), device=DEV) for a, s in zip(angs, SHAPES)]
    plan = batch.BatchPlan("A", xs, outs, angs)
    times = batch.autotune(plan, reps=2)
    assert set(times) == set(batch.FORWARD_VARIANTS)
    assert plan.variant == min(times, key=times.get)
    data = json.loads((tmp_path / "tune.json").read_text())
    assert any(v["best"] == plan.variant for v in data.values())
    batch._cache.clear()


def test_graph_captured_pair_replays_exactly():
    xs, _, angs = make_instances(5)
    ys = [torch.empty((a.numel(), s[4], s[2]), device=DEV) for a, s in zip(angs, SHAPES)]
    xb = [torch.empty_like(x) for x in xs]
    pa = batch.BatchPlan("A", xs, ys, angs, accumulate=False)
    pb = batch.BatchPlan("AT", ys, xb, angs, accumulate=False)
    pa.launch()
    pb.launch()
    torch.cuda.synchronize()
    ref_y = [t.clone() for t in ys]
    ref_x = [t.clone() for t in xb]
    graph = batch.capture_pair(pa, pb)
    for rep in range(3):
        for t in ys + xb:
            t.fill_(float("nan"))           # eager work between replays
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert all(torch.equal(a, b) for a, b in zip(ys, ref_y)), rep
        assert all(torch.equal(a, b) for a, b in zip(xb, ref_x)), rep


def test_plan_validation():
    x = torch.zeros((8, 8, 4), device=DEV)
    a = torch.zeros(5, device=DEV)
    with pytest.raises(ValueError):
        batch.BatchPlan("A", [x], [torch.zeros((6, 8, 4), device=DEV)], [a])
    with pytest.raises(ValueError):
        batch.BatchPlan("A", [x], [torch.zeros((5, 8, 3), device=DEV)], [a])
    plan = batch.BatchPlan("A", [x], [torch.zeros((5, 8, 4), device=DEV)], [a])
    with pytest.raises(RuntimeError, match="variant"):
        plan.launch(variant=3)


def test_grouped_plan_equals_plain_operator_bitwise():
    """Grouped launches give the plain results bit for bit: grouping uses the
    plain operator's variant by H, and where a single launch of a small
    volume drops from 2 slices per lane to 1, the two variants add the same
    terms in the same order.  The grouped A/AT pair also replays exactly from
    a graph."""
    xs, _, angs = make_instances(6)
    ys = [torch.empty((a.numel(), s[4], s[2]), device=DEV) for a, s in zip(angs, SHAPES)]
    xb = [torch.empty_like(x) for x in xs]
    pa = batch.GroupedPlan("A", xs, ys, angs)
    pb = batch.GroupedPlan("AT", ys, xb, angs)
    assert len(pa.plans) >= 2 and sum(pa.variants.values()) == len(SHAPES)
    pa.launch()
    pb.launch()
    torch.cuda.synchronize()
    for x, a, y, b, s in zip(xs, angs, ys, xb, SHAPES):
        assert torch.equal(y, ops.project(x, a, s[4]))
        assert torch.equal(b, ops.backproject(y, a, s[0], s[1]))
    ref_y = [t.clone() for t in ys]
    ref_x = [t.clone() for t in xb]
    graph = batch.capture_pair(pa, pb)
    for rep in range(2):
        for t in ys + xb:
            t.fill_(float("nan"))
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert all(torch.equal(u, v) for u, v in zip(ys, ref_y)), rep
        assert all(torch.equal(u, v) for u, v in zip(xb, ref_x)), rep


def test_concurrent_groups_equal_sequential():
    """Variant groups launched on parallel streams (and captured as parallel
    graph branches) give the sequential results bit for bit."""
    xs, _, angs = make_instances(8)
    ys = [torch.empty((a.numel(), s[4], s[2]), device=DEV) for a, s in zip(angs, SHAPES)]
    xb = [torch.empty_like(x) for x in xs]
    pa = batch.GroupedPlan("A", xs, ys, angs)
    pb = batch.GroupedPlan("AT", ys, xb, angs)
    pa.launch()
    pb.launch()
    torch.cuda.synchronize()
    ref_y = [t.clone() for t in ys]
    ref_x = [t.clone() for t in xb]
    for t in ys + xb:
        t.fill_(float("nan"))
    pa.launch(concurrent=True)
    pb.launch(concurrent=True)
    torch.cuda.synchronize()
    assert all(torch.equal(u, v) for u, v in zip(ys, ref_y))
    assert all(torch.equal(u, v) for u, v in zip(xb, ref_x))
    graph = batch.capture_pair(pa, pb, concurrent=True)
    for rep in range(3):
        for t in ys + xb:
            t.fill_(float("nan"))
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert all(torch.equal(u, v) for u, v in zip(ys, ref_y)), rep
        assert all(torch.equal(u, v) for u, v in zip(xb, ref_x)), rep


def test_narrow_forward_instances_form_their_own_group():
    """Forward instances with H <= 32 form their own group (variant 1, the
    single-instance operator's, so the grouped result equals it bit for
    bit); switching that group to variant 32 leaves the other groups alone
    and agrees with the plain operator to float32 rounding."""
    xs, _, angs = make_instances(9)
    ys = [torch.empty((a.numel(), s[4], s[2]), device=DEV) for a, s in zip(angs, SHAPES)]
    pa = batch.GroupedPlan("A", xs, ys, angs)
    narrow = [p for p in pa.plans if all(sh[2] <= 32 for sh in p.shapes)]
    assert len(narrow) == 1 and narrow[0].variant == 1
    assert sorted(sh[2] for sh in narrow[0].shapes) == sorted(s[2] for s in SHAPES if s[2] <= 32)
    assert all(all(sh[2] > 32 for sh in p.shapes) for p in pa.plans if p is not narrow[0])
    pa.launch()
    torch.cuda.synchronize()
    for x, a, y, s in zip(xs, angs, ys, SHAPES):
        assert torch.equal(y, ops.project(x, a, s[4]))
    ref = [t.clone() for t in ys]
    narrow[0].variant = 32
    for t in ys:
        t.fill_(float("nan"))
    pa.launch()
    torch.cuda.synchronize()
    for y, r, s in zip(ys, ref, SHAPES):
        if s[2] > 32 or max(s[0], s[1]) <= 128:
            assert torch.equal(y, r), s
        else:
            assert rel(y, r) <= 1e-6, s
