"""Container semantics of voxelproj_amd.forward / backward on the GPU.

Restates the reference's test/test_identity.py:21-110 (24 cases: same object
returned for device tensor / CPU tensor / numpy outputs; numpy in -> numpy
out, tensor in -> tensor on the input's device) and adds value checks:
results equal the FP64 oracle, and a prefilled output is accumulated into.
"""
import numpy as np
import pytest
import torch

import oracle as vpo
import voxelproj_amd as voxelproj

pytestmark = pytest.mark.gpu

SX, SY, H = 5, 6, 3
PX = SX + SY
n_angles = 7
y_shape = {0: (n_angles, PX, H), 2: (H, n_angles, PX)}
x_shape = {0: (SY, SX, H), 2: (H, SY, SX)}
angles = np.linspace(0, np.pi, n_angles).astype(np.float32)


@pytest.mark.parametrize("z_order", [0, 2])
def test_forward_keeps_id_for_cuda_tensor(z_order):
    x = torch.randn(x_shape[z_order], device="cuda")
    y = torch.zeros(y_shape[z_order], device="cuda")
    assert id(voxelproj.forward(x, angles, y, z_order=z_order)) == id(y)


@pytest.mark.parametrize("z_order", [0, 2])
def test_forward_keeps_id_for_cpu_tensor(z_order):
    x = torch.randn(x_shape[z_order])
    y = torch.zeros(y_shape[z_order])
    assert id(voxelproj.forward(x, angles, y, z_order=z_order)) == id(y)


@pytest.mark.parametrize("z_order", [0, 2])
def test_forward_keeps_id_for_numpy_array(z_order):
    x = np.random.randn(*x_shape[z_order]).astype(np.float32)
    y = np.zeros(y_shape[z_order], dtype=np.float32)
    assert id(voxelproj.forward(x, angles, y, z_order=z_order)) == id(y)


@pytest.mark.parametrize("x_type", ["torch_cpu", "torch_cuda", "numpy"])
@pytest.mark.parametrize("z_order", [0, 2])
def test_forward_returns_new_when_y_none(x_type, z_order):
    if x_type == "torch_cpu":
        x = torch.randn(x_shape[z_order])
    elif x_type == "torch_cuda":
        x = torch.randn(x_shape[z_order], device="cuda")
    else:
        x = np.random.randn(*x_shape[z_order]).astype(np.float32)
    y_out = voxelproj.forward(x, angles, None, z_order=z_order)
    if isinstance(x, np.ndarray):
        assert isinstance(y_out, np.ndarray)
    else:
        assert isinstance(y_out, torch.Tensor)
        assert y_out.device == x.device


@pytest.mark.parametrize("z_order", [0, 2])
def test_backward_keeps_id_for_cuda_tensor(z_order):
    x = torch.randn(x_shape[z_order], device="cuda")
    y = torch.zeros(y_shape[z_order], device="cuda")
    assert id(voxelproj.backward(y, angles, x, z_order=z_order)) == id(x)


@pytest.mark.parametrize("z_order", [0, 2])
def test_backward_keeps_id_for_cpu_tensor(z_order):
    x = torch.randn(x_shape[z_order])
    y = torch.zeros(y_shape[z_order])
    assert id(voxelproj.backward(y, angles, x, z_order=z_order)) == id(x)


@pytest.mark.parametrize("z_order", [0, 2])
def test_backward_keeps_id_for_numpy_array(z_order):
    x = np.random.randn(*x_shape[z_order]).astype(np.float32)
    y = np.zeros(y_shape[z_order], dtype=np.float32)
    assert id(voxelproj.backward(y, angles, x, z_order=z_order)) == id(x)


@pytest.mark.parametrize("y_type", ["torch_cpu", "torch_cuda", "numpy"])
@pytest.mark.parametrize("z_order", [0, 2])
def test_backward_returns_new_when_x_none(y_type, z_order):
    if y_type == "torch_cpu":
        y = torch.randn(y_shape[z_order])
    elif y_type == "torch_cuda":
        y = torch.randn(y_shape[z_order], device="cuda")
    else:
        y = np.random.randn(*y_shape[z_order]).astype(np.float32)
    x_out = voxelproj.backward(y, angles, None, z_order=z_order)
    if isinstance(y, np.ndarray):
        assert isinstance(x_out, np.ndarray)
    else:
        assert isinstance(x_out, torch.Tensor)
        assert x_out.device == y.device


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
    # test/test_adjoitness.py:20-51 through the public API (numpy in/out).
    for z_order in (0, 2):
        for PX_ in (SX + SY, min(SY, SX) - 1):
            ys = {0: (31, PX_, H), 2: (H, 31, PX_)}[z_order]
            ang = np.linspace(0, np.pi, 31).astype(np.float32)
            rng = np.random.default_rng(128437)
            x = rng.random(x_shape[z_order]).astype(np.float32)
            y = rng.random(ys).astype(np.float32)
            Ax = voxelproj.forward(x, ang, y=ys, z_order=z_order)
            ATy = voxelproj.backward(y, ang, x=x_shape[z_order], z_order=z_order)
            np.testing.assert_allclose(np.vdot(Ax, y), np.vdot(x, ATy), rtol=1e-5, atol=1e-5)


def test_non_contiguous_device_output_is_written_back():
    x = torch.randn(x_shape[0], device="cuda")
    y_big = torch.zeros((PX, n_angles, H), device="cuda")
    y_view = y_big.transpose(0, 1)  # (n_angles, PX, H), non-contiguous
    ref = voxelproj.forward(x, angles, (n_angles, PX, H))
    out = voxelproj.forward(x, angles, y_view)
    assert out is y_view
    torch.testing.assert_close(y_view, ref)
