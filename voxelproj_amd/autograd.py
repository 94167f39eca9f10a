"""Differentiable projector pair.

``Project`` applies A and uses the hand-written backprojection AT as its exact
backward (A is linear, so dL/dx = AT dL/dy); ``Backproject`` applies AT with
backward A.  Both run in float32 or float64 on the device; float64 is what
torch.autograd.gradcheck exercises.  Detector translations (du, dv) carry over
to the backward, which stays the exact adjoint.  The reference has no autograd support;
this follows the adjoint pairing of its two kernels
(voxelproj/csrc/include/sp_forward_z0.cuh, sp_backward_z0.cuh).
"""
import torch

from . import ops


class Project(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, angles, PX=None, z_order=0, du=None, dv=None):
        a = ops.device_angles(angles, x.device)
        SY, SX, _ = ops.vol_dims(x.shape, z_order)
        su, sv = ops.make_shifts(du, dv, a.numel(), x.device)
        ctx.save_for_backward(a)
        ctx.geom = (SY, SX, z_order, su, sv)
        return ops.project(x.contiguous(), a, PX, z_order, du=su, dv=sv)

    @staticmethod
    def backward(ctx, gy):
        (a,) = ctx.saved_tensors
        SY, SX, z_order, su, sv = ctx.geom
        gx = None
        if ctx.needs_input_grad[0]:
            gx = ops.backproject(gy.contiguous(), a, SY, SX, z_order, du=su, dv=sv)
        return gx, None, None, None, None, None


class Backproject(torch.autograd.Function):
    @staticmethod
    def forward(ctx, y, angles, SY=None, SX=None, z_order=0, du=None, dv=None):
        a = ops.device_angles(angles, y.device)
        _, PX, _ = ops.sino_dims(y.shape, z_order)
        su, sv = ops.make_shifts(du, dv, a.numel(), y.device)
        ctx.save_for_backward(a)
        ctx.geom = (PX, z_order, su, sv)
        return ops.backproject(y.contiguous(), a, SY, SX, z_order, du=su, dv=sv)

    @staticmethod
    def backward(ctx, gx):
        (a,) = ctx.saved_tensors
        PX, z_order, su, sv = ctx.geom
        gy = None
        if ctx.needs_input_grad[0]:
            gy = ops.project(gx.contiguous(), a, PX, z_order, du=su, dv=sv)
        return gy, None, None, None, None, None, None


def project(x, angles, PX=None, z_order=0, du=None, dv=None):
    """Differentiable A x."""
    return Project.apply(x, angles, PX, z_order, du, dv)


def backproject(y, angles, SY=None, SX=None, z_order=0, du=None, dv=None):
    """Differentiable AT y."""
    return Backproject.apply(y, angles, SY, SX, z_order, du, dv)


def least_squares_gradient(x, b, angles, PX=None, z_order=0, du=None, dv=None):
    """Loss 1/2 ||A x - b||^2 and its gradient AT (A x - b), without autograd."""
    a = ops.device_angles(angles, x.device)
    su, sv = ops.make_shifts(du, dv, a.numel(), x.device)
    r = ops.project(x, a, PX, z_order, du=su, dv=sv)
    r -= b
    loss = 0.5 * torch.dot(r.flatten().double(), r.flatten().double())
    SY, SX, _ = ops.vol_dims(x.shape, z_order)
    g = ops.backproject(r, a, SY, SX, z_order, du=su, dv=sv)
    return loss, g
