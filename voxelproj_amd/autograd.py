"""Differentiable projector pair.

``Project`` applies A and uses the hand-written backprojection AT as its exact
backward (A is linear, so dL/dx = AT dL/dy); ``Backproject`` applies AT with
backward A.  Both run in float32 or float64 on the device; float64 is what
torch.autograd.gradcheck exercises.  The reference has no autograd support;
this follows the adjoint pairing of its two kernels
(voxelproj/csrc/include/sp_forward_z0.cuh, sp_backward_z0.cuh).
"""
import torch

from . import ops


class Project(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, angles, PX=None, z_order=0):
        a = ops.device_angles(angles, x.device)
        SY, SX, _ = ops.vol_dims(x.shape, z_order)
        ctx.save_for_backward(a)
        ctx.geom = (SY, SX, z_order)
        return ops.project(x.contiguous(), a, PX, z_order)

    @staticmethod
    def backward(ctx, gy):
        (a,) = ctx.saved_tensors
        SY, SX, z_order = ctx.geom
        gx = None
        if ctx.needs_input_grad[0]:
            gx = ops.backproject(gy.contiguous(), a, SY, SX, z_order)
        return gx, None, None, None


class Backproject(torch.autograd.Function):
    @staticmethod
    def forward(ctx, y, angles, SY=None, SX=None, z_order=0):
        a = ops.device_angles(angles, y.device)
        _, PX, _ = ops.sino_dims(y.shape, z_order)
        ctx.save_for_backward(a)
        ctx.geom = (PX, z_order)
        return ops.backproject(y.contiguous(), a, SY, SX, z_order)

    @staticmethod
    def backward(ctx, gx):
        (a,) = ctx.saved_tensors
        PX, z_order = ctx.geom
        gy = None
        if ctx.needs_input_grad[0]:
            gy = ops.project(gx.contiguous(), a, PX, z_order)
        return gy, None, None, None, None


def project(x, angles, PX=None, z_order=0):
    """Differentiable A x."""
    return Project.apply(x, angles, PX, z_order)


def backproject(y, angles, SY=None, SX=None, z_order=0):
    """Differentiable AT y."""
    return Backproject.apply(y, angles, SY, SX, z_order)


def least_squares_gradient(x, b, angles, PX=None, z_order=0):
    """Loss 1/2 ||A x - b||^2 and its gradient AT (A x - b), without autograd."""
    r = ops.project(x, angles, PX, z_order)
    r -= b
    loss = 0.5 * torch.dot(r.flatten().double(), r.flatten().double())
    SY, SX, _ = ops.vol_dims(x.shape, z_order)
    g = ops.backproject(r, angles, SY, SX, z_order)
    return loss, g
