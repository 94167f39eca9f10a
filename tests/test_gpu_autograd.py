"""Differentiation and solvers on the GPU.

* FP64 kernels against the FP64 oracle (same operator, both in double);
* torch.autograd.gradcheck of Project / Backproject in FP64 on a 24x24x8
  volume with 30 angles, central finite differences with h = 1e-6;
* adjoint identity of the FP64 pair;
* least squares 1/2||Ax - b||^2: autograd gradient == AT(Ax - b);
* SIRT: residual decreases monotonically on an ellipsoid phantom.
"""
import numpy as np
import pytest
import torch

import oracle as vpo
from voxelproj_amd import autograd, ops, recon

pytestmark = pytest.mark.gpu
DEV = "cuda"


@pytest.mark.parametrize("z", [0, 2])
def test_f64_kernels_match_oracle(z):
    SY, SX, H, A, PX = 24, 20, 8, 30, 27
    ang = np.linspace(0, np.pi, A, endpoint=False).astype(np.float32)
    g = torch.Generator().manual_seed(4)
    x = torch.randn(ops.vol_shape(SY, SX, H, z), generator=g, dtype=torch.float64)
    yv = torch.randn(ops.sino_shape(A, PX, H, z), generator=g, dtype=torch.float64)
    yg = ops.project(x.to(DEV), ang, PX, z).cpu().numpy()
    xg = ops.backproject(yv.to(DEV), ang, SY, SX, z).cpu().numpy()
    yo = vpo.forward_f64(x.numpy(), ang, PX, z)
    xo = vpo.backward_f64(yv.numpy(), ang, SY, SX, z)
    assert np.abs(yg - yo).max() <= 1e-11 * np.abs(yo).max()
    assert np.abs(xg - xo).max() <= 1e-11 * np.abs(xo).max()


def test_gradcheck_project_f64():
    SY = SX = 24
    H, A = 8, 30
    ang = torch.tensor(np.linspace(0, np.pi, A, endpoint=False).astype(np.float32), device=DEV)
    g = torch.Generator(device=DEV).manual_seed(1)
    x = torch.randn((SY, SX, H), device=DEV, dtype=torch.float64, generator=g, requires_grad=True)
    assert torch.autograd.gradcheck(lambda t: autograd.project(t, ang, SX, 0), (x,),
                                    eps=1e-6, atol=1e-6, rtol=1e-5)


def test_gradcheck_backproject_f64():
    SY = SX = 24
    H, A = 8, 30
    ang = torch.tensor(np.linspace(0, np.pi, A, endpoint=False).astype(np.float32), device=DEV)
    g = torch.Generator(device=DEV).manual_seed(2)
    y = torch.randn((H, A, SX), device=DEV, dtype=torch.float64, generator=g, requires_grad=True)
    assert torch.autograd.gradcheck(lambda t: autograd.backproject(t, ang, SY, SX, 2), (y,),
                                    eps=1e-6, atol=1e-6, rtol=1e-5)


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-13), (torch.float32, 1e-6)])
def test_adjoint_identity(dtype, tol):
    SY, SX, H, A, PX = 40, 36, 16, 45, 60
    ang = torch.tensor(np.linspace(0, np.pi, A).astype(np.float32), device=DEV)
    g = torch.Generator(device=DEV).manual_seed(3)
    x = torch.randn((SY, SX, H), device=DEV, dtype=dtype, generator=g)
    y = torch.randn((A, PX, H), device=DEV, dtype=dtype, generator=g)
    Ax = ops.project(x, ang, PX)
    ATy = ops.backproject(y, ang, SY, SX)
    lhs = torch.dot(Ax.double().flatten(), y.double().flatten())
    rhs = torch.dot(x.double().flatten(), ATy.double().flatten())
    gap = abs(float(lhs - rhs)) / float(Ax.double().norm() * y.double().norm())
    assert gap <= tol


def test_least_squares_gradient_matches_autograd():
    SY = SX = 32
    H, A = 8, 40
    ang = torch.tensor(np.linspace(0, np.pi, A, endpoint=False).astype(np.float32), device=DEV)
    x_true = recon.ellipsoid_phantom(SY, SX, H, seed=3, device=DEV)
    b = ops.project(x_true, ang)
    x = torch.zeros_like(x_true, requires_grad=True)
    loss = 0.5 * (autograd.project(x, ang) - b).pow(2).sum()
    loss.backward()
    _, g = autograd.least_squares_gradient(x.detach().clone(), b, ang)
    assert torch.allclose(x.grad, g, rtol=1e-5, atol=1e-5 * float(g.abs().max()))


def test_sirt_residual_decreases():
    SY = SX = 64
    H, A = 16, 90
    ang = np.linspace(0, np.pi, A, endpoint=False).astype(np.float32)
    x_true = recon.ellipsoid_phantom(SY, SX, H, seed=5, device=DEV)
    solver = recon.SIRT(ang, SY, SX, H, device=DEV)
    b = ops.project(x_true, solver.angles)
    x = torch.zeros_like(x_true)
    res = [solver.residual(x, b)]
    for _ in range(10):
        solver.iterate(x, b, 1)
        res.append(solver.residual(x, b))
    assert all(r1 < r0 for r0, r1 in zip(res, res[1:]))
    assert res[-1] < 0.2 * res[0]


def test_gradient_descent_converges():
    SY = SX = 48
    H, A = 8, 60
    ang = torch.tensor(np.linspace(0, np.pi, A, endpoint=False).astype(np.float32), device=DEV)
    x_true = recon.ellipsoid_phantom(SY, SX, H, seed=6, device=DEV)
    b = ops.project(x_true, ang)
    step = recon.lipschitz_step(ang, SY, SX, H, device=DEV)
    x = torch.zeros_like(x_true)
    losses = recon.gradient_descent(x, b, ang, 30, step)
    assert losses[-1] < 0.05 * losses[0]
    assert all(l1 <= l0 * (1 + 1e-6) for l0, l1 in zip(losses, losses[1:]))
