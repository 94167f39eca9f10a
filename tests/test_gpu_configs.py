"""Launch configurations of the single-instance operators.

A single launch picks its kernel configuration from the volume and the GPU.
A: 1 or 2 slices per lane, and up to 128 voxels per side, for volumes whose
grid would leave CUs idle, the narrow forward strips (16 pixels).  None of
these choices changes the order in which an output adds its terms, so every
A configuration with 1 or 2 slices per lane must give the plain operator's
result bit for bit.  The batched launches with an explicit variant use the
default configurations (28-pixel strips), so comparing them with the single
launch checks exactly that.  Above 128 slices a single launch of A takes
the strip width (22, 20, 18 or 16 pixels) whose grid fills the device best,
and equals the batched launch of that variant bit for bit.
AT: 4 slices per lane, 256 slices per block above 128 slices, below that
rows of 8, 16 or 32 lanes per voxel (blocks of 32, 64 or 128 slices); all of
them add each voxel's terms in the same order, so the single launch equals
the batched launches of variants 4, 32, 64 and 128 bit for bit, and the 1-
and 2-slice variants (another order) agree with it to float32 rounding.
"""
import numpy as np
import pytest
import torch

import oracle as vpo
from voxelproj_amd import batch, ops

pytestmark = pytest.mark.gpu
DEV = "cuda"

# (SY, SX, H, n_angles, PX): tiny (small configurations), small volumes whose
# default grid has fewer workgroups than CUs, and mid-size volumes on the
# default configurations; ragged H and truncated detectors included.
SHAPES = [(24, 20, 9, 30, 32), (96, 96, 64, 120, 96), (128, 128, 128, 192, 128),
          (72, 200, 37, 77, 150), (256, 256, 40, 360, 256), (160, 160, 100, 192, 120)]


def _inputs(shape, seed):
    SY, SX, H, A, PX = shape
    g = torch.Generator(device=DEV).manual_seed(seed)
    x = torch.randn((SY, SX, H), device=DEV, generator=g)
    y = torch.randn((A, PX, H), device=DEV, generator=g)
    rng = np.random.default_rng(seed)
    ang = torch.tensor(np.sort(rng.uniform(0, np.pi, A)).astype(np.float32), device=DEV)
    return x, y, ang


@pytest.mark.parametrize("shape", SHAPES)
def test_single_launch_equals_every_configuration_of_its_family(shape):
    SY, SX, H, A, PX = shape
    x, y, ang = _inputs(shape, 17)
    print(f"shape {shape}", flush=True)
    ya = ops.project(x, ang, PX)
    xb = ops.backproject(y, ang, SY, SX)
    for v in (1, 2):
        yv = torch.empty_like(ya)
        batch.BatchPlan("A", [x], [yv], [ang], accumulate=False, variant=v).launch()
        assert torch.equal(yv, ya), ("A", v)
    outs = {}
    for v in batch.BACKWARD_VARIANTS:
        xv = torch.empty_like(xb)
        batch.BatchPlan("AT", [y], [xv], [ang], accumulate=False, variant=v).launch()
        outs[v] = xv
    for v in (4, 32, 64, 128):
        assert torch.equal(outs[v], xb), ("AT", v)
    assert torch.equal(outs[1], outs[2])
    assert float((outs[1] - xb).norm() / xb.norm()) <= 1e-6


@pytest.mark.parametrize("shape", SHAPES[:3])
def test_small_configurations_match_oracle(shape):
    SY, SX, H, A, PX = shape
    x, y, ang = _inputs(shape, 5)
    hs = slice(0, min(H, 6))  # slices are independent: check a few against FP64
    xs = x[:, :, hs].contiguous()
    ys = y[:, :, hs].contiguous()
    a = ang.cpu().numpy()
    got_y = ops.project(x, ang, PX)[:, :, hs].cpu().numpy()
    got_x = ops.backproject(y, ang, SY, SX)[:, :, hs].cpu().numpy()
    ref_y = vpo.forward_f64(xs.cpu().numpy(), a, PX, 0)
    ref_x = vpo.backward_f64(ys.cpu().numpy(), a, SY, SX, 0)
    for got, ref in ((got_y, ref_y), (got_x, ref_x)):
        assert np.linalg.norm(got - ref) / np.linalg.norm(ref) <= 1e-6
        assert np.abs(got - ref).max() / np.abs(ref).max() <= 1e-5


# Single launches of A above 128 slices pick the strip width (22, 20, 18 or
# 16 pixels: variants 4, 420, 418, 416) whose grid fills the device best;
# these sizes reach each of them on a 256-CU MI355X.
STRIP_SHAPES = [(256, 256, 256, 192, 256), (288, 288, 288, 192, 288), (320, 320, 320, 192, 320),
                (160, 200, 300, 96, 230), (384, 64, 260, 40, 390), (256, 256, 300, 720, 256)]


def test_single_launch_strip_width_equals_its_batched_variant():
    """A single launch equals the batched launch of the variant
    vp_launch_variant names for its geometry bit for bit, and agrees with the
    other strip widths to float32 rounding; a few slices against the FP64
    oracle."""
    from voxelproj_amd import _lib
    lib = _lib.load()
    seen = set()
    for shape in STRIP_SHAPES:
        SY, SX, H, A, PX = shape
        print(f"strip width {shape}", flush=True)
        x, _, ang = _inputs(shape, 41)
        v = int(lib.vp_launch_variant(0, A, SY, SX, H, PX))
        assert v in (4, 416, 418, 420), v
        seen.add(v)
        ya = ops.project(x, ang, PX)
        yv = torch.empty_like(ya)
        batch.BatchPlan("A", [x], [yv], [ang], accumulate=False, variant=v).launch()
        assert torch.equal(yv, ya), (shape, v)
        for w in (4, 416, 418, 420):
            if w != v:
                yw = torch.empty_like(ya)
                batch.BatchPlan("A", [x], [yw], [ang], accumulate=False, variant=w).launch()
                assert float((yw.double() - ya.double()).norm() / ya.double().norm()) <= 1e-6, (shape, w)
        hs = slice(H - 3, H)
        ref = vpo.forward_f64(x[:, :, hs].contiguous().cpu().numpy(), ang.cpu().numpy(), PX, 0)
        got = ya[:, :, hs].cpu().numpy().astype(np.float64)
        assert np.linalg.norm(got - ref) / np.linalg.norm(ref) <= 1e-6, shape
    if torch.cuda.get_device_properties(0).multi_processor_count == 256:
        assert seen == {4, 416, 418, 420}, seen


# Forward variant 32 (1 slice per lane, two angles per sweep wave, 32
# slices per block): H <= 32, H > 32 over several slice blocks, ragged H,
# one voxel, and volumes over 128 voxels per side (lines split into pieces).
TWO_ANGLE_SHAPES = [(24, 20, 9, 30, 32), (96, 96, 32, 120, 96), (128, 128, 32, 192, 128),
                    (100, 60, 70, 50, 110), (1, 1, 16, 7, 3), (33, 41, 30, 37, 45),
                    (72, 200, 31, 77, 150), (256, 256, 24, 360, 256)]


@pytest.mark.parametrize("shape", TWO_ANGLE_SHAPES)
def test_forward_two_angles_per_wave(shape):
    """Variant 32 against variant 1: bit for bit where no line is split
    (at most 128 voxels per side), to float32 rounding elsewhere; a few
    slices against the FP64 oracle, with accumulation into a nonzero
    sinogram."""
    SY, SX, H, A, PX = shape
    print(f"two angles per wave {shape}", flush=True)
    x, y0, ang = _inputs(shape, 29)
    ang = torch.tensor(np.random.default_rng(A).uniform(0, 2 * np.pi, A).astype(np.float32), device=DEV)
    y1, y32 = y0.clone(), y0.clone()
    batch.BatchPlan("A", [x], [y1], [ang], accumulate=True, variant=1).launch()
    batch.BatchPlan("A", [x], [y32], [ang], accumulate=True, variant=32).launch()
    if max(SY, SX) <= 128:
        assert torch.equal(y32, y1)
    else:
        d = (y32 - y0).double()
        r = (y1 - y0).double()
        assert float((d - r).norm() / r.norm()) <= 1e-6
    hs = slice(0, min(H, 4))
    ref = vpo.forward_f64(x[:, :, hs].contiguous().cpu().numpy(), ang.cpu().numpy(), PX, 0)
    got = (y32 - y0)[:, :, hs].cpu().numpy().astype(np.float64)
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) <= 1e-6


# x-dominant angles with cos > 0 and cos < 0 and y-dominant angles with
# sin > 0 and sin < 0, interleaved, so every 8- or 12-angle block of the
# forward kernel holds, in both passes, waves whose centre pixel runs
# forwards and backwards along the staged lines.  The block stages each line
# once for all its waves.
MIXED = np.array([0.05, np.pi - 0.05, 1.2, -1.2, 0.3, np.pi - 0.3, 1.0, np.pi + 1.0,
                  0.01, np.pi - 0.01, 1.5, -1.5, 0.4, np.pi - 0.45, 0.9, np.pi + 0.9],
                 dtype=np.float32)


@pytest.mark.parametrize("shape", [(60, 52, 256, 70), (41, 47, 130, 40), (33, 30, 7, 50)],
                         ids=["vec4", "ragged", "element"])
def test_blocks_with_both_line_directions(shape):
    SY, SX, H, PX = shape
    g = torch.Generator(device=DEV).manual_seed(23)
    x = torch.randn((SY, SX, H), device=DEV, generator=g)
    ang = torch.tensor(MIXED, device=DEV)
    hs = slice(0, 5)
    ref = vpo.forward_f64(x[:, :, hs].contiguous().cpu().numpy(), MIXED, PX, 0)
    outs = {"single": ops.project(x, ang, PX)}
    for v in batch.FORWARD_VARIANTS:
        yv = torch.empty_like(outs["single"])
        batch.BatchPlan("A", [x], [yv], [ang], accumulate=False, variant=v).launch()
        outs[v] = yv
    for k, y in outs.items():
        got = y.cpu().numpy()
        assert np.linalg.norm(got[:, :, hs] - ref) / np.linalg.norm(ref) <= 1e-6, k
        assert np.abs(got[:, :, hs] - ref).max() / np.abs(ref).max() <= 1e-5, k
        # every slice sees the same footprints: the checked slices stand for all
        # (the last ones come from other lanes and, for H > 64, other blocks)
        tail = slice(H - 3, H)
        ref_t = vpo.forward_f64(x[:, :, tail].contiguous().cpu().numpy(), MIXED, PX, 0)
        assert np.linalg.norm(got[:, :, tail] - ref_t) / np.linalg.norm(ref_t) <= 1e-6, k


# H above 256 with a last block of at most 128 slices: single launches of AT
# give that block to a row kernel of its size.
RAGGED_AT_SHAPES = [(40, 36, 257, 45, 52), (33, 41, 288, 29, 50), (24, 30, 300, 60, 40),
                    (20, 20, 320, 17, 31), (16, 24, 384, 23, 26), (9, 13, 600, 11, 20), (1, 1, 259, 5, 3)]


@pytest.mark.parametrize("shape", RAGGED_AT_SHAPES)
def test_at_ragged_tail_equals_one_kernel(shape):
    """The split launch (4-slice kernel on whole 256-slice blocks, row kernel
    on the rest) equals the 4-slice kernel over all of H bit for bit, with
    accumulation, and the FP64 oracle on the last slices."""
    SY, SX, H, A, PX = shape
    print(f"ragged AT {shape}", flush=True)
    _, y, ang = _inputs(shape, 53)
    x0 = torch.randn((SY, SX, H), device=DEV, generator=torch.Generator(device=DEV).manual_seed(H))
    got = ops.backproject(y, ang, SY, SX, 0, out=x0.clone())
    one = x0.clone()
    batch.BatchPlan("AT", [y], [one], [ang], accumulate=True, variant=4).launch()
    assert torch.equal(got, one)
    hs = slice(H - 4, H)
    ref = vpo.backward_f64(y[:, :, hs].contiguous().cpu().numpy().astype(np.float64), ang.cpu().numpy(), SY, SX, 0)
    sums = (got - x0)[:, :, hs].cpu().numpy().astype(np.float64)
    scale = np.linalg.norm(ref)
    assert np.linalg.norm(sums - ref) / scale <= 1e-6 + 2e-7 * np.linalg.norm(x0[:, :, hs].cpu().numpy()) / scale


# H above 256 with a last block of at most 128 slices: single launches of A
# give that block to a 1- or 2-slice kernel of the same strips and pieces.
RAGGED_A_SHAPES = [(40, 36, 257, 45, 52), (33, 41, 288, 29, 50), (24, 30, 300, 60, 40),
                   (20, 20, 320, 17, 31), (16, 24, 384, 23, 26), (9, 13, 600, 11, 20), (1, 1, 259, 5, 3),
                   (256, 256, 288, 192, 256)]


@pytest.mark.parametrize("shape", RAGGED_A_SHAPES)
def test_a_ragged_tail_equals_one_kernel(shape):
    """The split launch (4-slice kernel on whole 256-slice blocks, a 1- or
    2-slice kernel of the same strips on the rest) equals the 4-slice kernel
    of the same strip width over all of H bit for bit, with accumulation, and
    the FP64 oracle on the last slices."""
    from voxelproj_amd import _lib
    SY, SX, H, A, PX = shape
    print(f"ragged A {shape}", flush=True)
    x, _, ang = _inputs(shape, 59)
    y0 = torch.randn((A, PX, H), device=DEV, generator=torch.Generator(device=DEV).manual_seed(H))
    got = ops.project(x, ang, PX, 0, out=y0.clone())
    v = int(_lib.load().vp_launch_variant(0, A, SY, SX, H, PX))
    one = y0.clone()
    batch.BatchPlan("A", [x], [one], [ang], accumulate=True, variant=v).launch()
    assert torch.equal(got, one), v
    hs = slice(H - 4, H)
    ref = vpo.forward_f64(x[:, :, hs].contiguous().cpu().numpy(), ang.cpu().numpy(), PX, 0)
    sums = (got - y0)[:, :, hs].cpu().numpy().astype(np.float64)
    scale = np.linalg.norm(ref)
    assert np.linalg.norm(sums - ref) / scale <= 1e-6 + 2e-7 * np.linalg.norm(y0[:, :, hs].cpu().numpy()) / scale
