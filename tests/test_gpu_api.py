"""Container semantics of voxelproj_amd.forward / backward on the GPU.

The container contract of voxelproj/wrappers.py:28-122 (what
test/test_identity.py:21-110 exercises): a given output container comes back
as the same object, and without one the result follows the input's kind.
Written here as one table over op x container kind x layout (24 cases), plus
value checks: results equal the FP64 oracle, and a prefilled output is
accumulated into.
"""
import numpy as np
import pytest
import torch

import oracle as vpo
import voxelproj_amd as voxelproj

pytestmark = pytest.mark.gpu

# Geometry of the container cases: a non-square volume, a detector wider than
# the volume, an odd angle count.
SY, SX, H = 7, 4, 5
PX = 9
n_angles = 9
angles = np.linspace(0, np.pi, n_angles, endpoint=False).astype(np.float32)


def vol_shape(z):
    return (SY, SX, H) if z == 0 else (H, SY, SX)


def sino_shape(z):
    return (n_angles, PX, H) if z == 0 else (H, n_angles, PX)


def make(kind, shape, fill):
    """A container of `kind` ("device", "host" tensor or "numpy") holding `fill` data."""
    a = fill(shape)
    if kind == "numpy":
        return a
    t = torch.from_numpy(a)
    return t.cuda() if kind == "device" else t


_rng = np.random.default_rng(3)


def _randn(shape):
    return _rng.standard_normal(shape).astype(np.float32)


def _zeros(shape):
    return np.zeros(shape, dtype=np.float32)


# The contract (voxelproj/wrappers.py:28-122): a container passed as the output
# is accumulated into and returned as the same object, whatever its kind; with
# output None the result is numpy for numpy input and otherwise a tensor on the
# input's device.  One row per op x input kind x output kind; each row runs in
# both layouts.
KINDS = ("device", "host", "numpy")
OUTPUT_ROWS = [(op, kind) for op in ("forward", "backward") for kind in KINDS]
NONE_ROWS = [(op, kind) for op in ("forward", "backward") for kind in KINDS]


def _call(op, inp, out, z):
    if op == "forward":
        return voxelproj.forward(inp, angles, out, z_order=z)
    return voxelproj.backward(inp, angles, out, z_order=z)


def _shapes(op, z):
    return (vol_shape(z), sino_shape(z)) if op == "forward" else (sino_shape(z), vol_shape(z))


@pytest.mark.parametrize("z", [0, 2])
@pytest.mark.parametrize("op,kind", OUTPUT_ROWS)
def test_output_container_is_returned(op, kind, z):
    in_shape, out_shape = _shapes(op, z)
    inp = make(kind, in_shape, _randn)
    out = make(kind, out_shape, _zeros)
    assert _call(op, inp, out, z) is out


@pytest.mark.parametrize("z", [0, 2])
@pytest.mark.parametrize("op,kind", NONE_ROWS)
def test_result_kind_follows_input(op, kind, z):
    in_shape, out_shape = _shapes(op, z)
    inp = make(kind, in_shape, _randn)
    res = _call(op, inp, None, z)
    if kind == "numpy":
        assert isinstance(res, np.ndarray)
    else:
        assert isinstance(res, torch.Tensor) and res.device == inp.device
    # default shapes (wrappers.py:40-43, 88-91): the detector spans max(SY, SX)
    # pixels; a backprojected volume is PX x PX
    if op == "forward":
        want = (n_angles, max(SY, SX), H) if z == 0 else (H, n_angles, max(SY, SX))
    else:
        want = (PX, PX, H) if z == 0 else (H, PX, PX)
    assert tuple(res.shape) == want


# ------------------------------------------------------------- values
def _to_z(a, z):
    return a if z == 0 else np.ascontiguousarray(np.transpose(a, (2, 0, 1)))


@pytest.mark.parametrize("kind", ["numpy", "torch_cpu", "torch_cuda"])
@pytest.mark.parametrize("z_order", [0, 2])
def test_values_and_accumulation(kind, z_order):
    rng = np.random.default_rng(10)
    x0 = rng.standard_normal((SY, SX, H)).astype(np.float32)
    y0 = rng.standard_normal((n_angles, PX, H)).astype(np.float32)
    Ax = _to_z(vpo.forward_f64(x0, angles, PX, 0), z_order)
    ATy = _to_z(vpo.backward_f64(y0, angles, SY, SX, 0), z_order)
    x, yv = _to_z(x0, z_order), _to_z(y0, z_order)
    prior_y = rng.standard_normal(Ax.shape).astype(np.float32)
    prior_x = rng.standard_normal(ATy.shape).astype(np.float32)

    def box(a):
        if kind == "numpy":
            return a.copy()
        t = torch.tensor(a)
        return t.cuda() if kind == "torch_cuda" else t

    def unbox(a):
        return a if isinstance(a, np.ndarray) else a.cpu().numpy()

    out = unbox(voxelproj.forward(box(x), angles, Ax.shape, z_order=z_order))
    np.testing.assert_allclose(out, Ax, rtol=0, atol=2e-6 * np.abs(Ax).max())
    acc = unbox(voxelproj.forward(box(x), angles, box(prior_y), z_order=z_order))
    np.testing.assert_allclose(acc, prior_y + Ax, rtol=0, atol=4e-6 * np.abs(prior_y + Ax).max())
    outb = unbox(voxelproj.backward(box(yv), angles, x.shape, z_order=z_order))
    np.testing.assert_allclose(outb, ATy, rtol=0, atol=2e-6 * np.abs(ATy).max())
    accb = unbox(voxelproj.backward(box(yv), angles, box(prior_x), z_order=z_order))
    np.testing.assert_allclose(accb, prior_x + ATy, rtol=0, atol=4e-6 * np.abs(prior_x + ATy).max())


def test_reference_adjointness_through_api():
    # The adjointness geometries of test/test_adjoitness.py:12-51 (5 x 6 x 3
    # volume, 31 angles, a detector covering the diagonal and a truncated one)
    # through the public API (numpy in/out), at a tighter tolerance.
    sy, sx, h, na = 6, 5, 3, 31
    for z_order in (0, 2):
        for px in (sx + sy, min(sy, sx) - 1):
            xs = (sy, sx, h) if z_order == 0 else (h, sy, sx)
            ys = (na, px, h) if z_order == 0 else (h, na, px)
            ang = np.linspace(0, np.pi, na).astype(np.float32)
            rng = np.random.default_rng(128437)
            x = rng.random(xs).astype(np.float32)
            y = rng.random(ys).astype(np.float32)
            Ax = voxelproj.forward(x, ang, y=ys, z_order=z_order)
            ATy = voxelproj.backward(y, ang, x=xs, z_order=z_order)
            np.testing.assert_allclose(np.vdot(Ax, y), np.vdot(x, ATy), rtol=1e-5, atol=1e-5)


def test_reference_import_names():
    # Caller code written for the reference imports `voxelproj` and its
    # compiled module `voxelproj._CU` (voxelproj/__init__.py:1-5,
    # voxelproj/wrappers.py:4-5); both resolve to this library.
    import voxelproj as vp_ref_name
    from voxelproj._CU import backward_z0, forward_z0, forward_z2

    rng = np.random.default_rng(4)
    x = rng.standard_normal((SY, SX, H)).astype(np.float32)
    want = vpo.forward_f64(x, angles, max(SY, SX), 0)
    got = vp_ref_name.forward(x, angles, None)
    assert isinstance(got, np.ndarray)
    assert np.linalg.norm(got - want) <= 1e-6 * np.linalg.norm(want)
    xd = torch.from_numpy(x).cuda()
    yd = torch.zeros(want.shape, device="cuda")
    ad = torch.from_numpy(angles).cuda()
    forward_z0(xd, yd, ad)
    assert np.linalg.norm(yd.cpu().numpy() - want) <= 1e-6 * np.linalg.norm(want)
    y2 = torch.zeros((H, n_angles, max(SY, SX)), device="cuda")
    forward_z2(xd.permute(2, 0, 1).contiguous(), y2, ad)
    torch.testing.assert_close(y2.permute(1, 2, 0), yd, rtol=0, atol=1e-5)
    xb = torch.zeros((SY, SX, H), device="cuda")
    backward_z0(yd, xb, ad)
    wantb = vpo.backward_f64(yd.cpu().numpy(), angles, SY, SX, 0)
    assert np.linalg.norm(xb.cpu().numpy() - wantb) <= 1e-6 * np.linalg.norm(wantb)


def test_non_contiguous_device_output_is_written_back():
    x = torch.randn(vol_shape(0), device="cuda")
    y_big = torch.zeros((PX, n_angles, H), device="cuda")
    y_view = y_big.transpose(0, 1)  # (n_angles, PX, H), non-contiguous
    ref = voxelproj.forward(x, angles, (n_angles, PX, H))
    out = voxelproj.forward(x, angles, y_view)
    assert out is y_view
    torch.testing.assert_close(y_view, ref)
