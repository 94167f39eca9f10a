"""Reconstruction helpers on the device: phantoms, SIRT, gradient descent.

Not part of the reference (which ships only the operator pair); they are the
callers of A/AT that the batched and served forms run.

SIRT follows the textbook form (Gregor & Benson, IEEE TMI 27(7), 2008):
    x <- x + C * AT( R * (b - A x) ),  R = 1 / (A 1),  C = 1 / (AT 1),
with R and C set to 0 where the row / column sums vanish.
"""
import math

import torch

from . import ops


def ellipsoid_phantom(SY, SX, H, n_ellipsoids=None, seed=0, z_order=0, device="cuda",
                      dtype=torch.float32):
    """Piecewise-constant phantom: sum of 8-16 random ellipsoids, intensities in [0, 1].

    Coordinates are normalised to [-1, 1] on every axis; each ellipsoid has a
    random centre, semi-axes and in-plane rotation.  Deterministic in `seed`.
    """
    g = torch.Generator().manual_seed(seed)
    n = int(torch.randint(8, 17, (1,), generator=g)) if n_ellipsoids is None else n_ellipsoids
    cen = (torch.rand(n, 3, generator=g) * 1.2 - 0.6).tolist()
    ax = (torch.rand(n, 3, generator=g) * 0.35 + 0.05).tolist()
    rot = (torch.rand(n, generator=g) * math.pi).tolist()
    val = torch.rand(n, generator=g).tolist()
    zz = torch.linspace(-1, 1, H, device=device, dtype=torch.float32)
    yy = torch.linspace(-1, 1, SY, device=device, dtype=torch.float32)
    xx = torch.linspace(-1, 1, SX, device=device, dtype=torch.float32)
    Yg, Xg, Zg = torch.meshgrid(yy, xx, zz, indexing="ij")  # (SY, SX, H)
    vol = torch.zeros((SY, SX, H), device=device, dtype=torch.float32)
    for i in range(n):
        c, s = math.cos(rot[i]), math.sin(rot[i])
        dx, dy, dz = Xg - cen[i][0], Yg - cen[i][1], Zg - cen[i][2]
        u = (c * dx + s * dy) / ax[i][0]
        v = (-s * dx + c * dy) / ax[i][1]
        w = dz / ax[i][2]
        vol += val[i] * ((u * u + v * v + w * w) <= 1.0).to(torch.float32)
    if z_order == 2:
        vol = vol.permute(2, 0, 1)
    return vol.contiguous().to(dtype)


class SIRT:
    """SIRT solver with device-resident weights and work buffers.

    All buffers are allocated once; `iterate` then performs only the
    projector launches and elementwise updates on the current stream.
    """

    def __init__(self, angles, SY, SX, H, PX=None, z_order=0, device="cuda",
                 dtype=torch.float32, du=None, dv=None):
        self.z = z_order
        self.SY, self.SX, self.H = SY, SX, H
        self.PX = max(SY, SX) if PX is None else PX
        self.angles = ops.device_angles(angles, device)
        A = self.angles.numel()
        # optional detector translations, fixed for the solver's lifetime
        self.du, self.dv = ops.make_shifts(du, dv, A, self.angles.device)
        self.vshape = ops.vol_shape(SY, SX, H, z_order)
        self.sshape = ops.sino_shape(A, self.PX, H, z_order)
        ones_v = torch.ones(self.vshape, device=device, dtype=dtype)
        ones_s = torch.ones(self.sshape, device=device, dtype=dtype)
        rs = ops.project(ones_v, self.angles, self.PX, z_order, du=self.du, dv=self.dv)
        cs = ops.backproject(ones_s, self.angles, SY, SX, z_order, du=self.du, dv=self.dv)
        self.R = rs.clamp_min(1e-6).reciprocal_().mul_(rs > 1e-6)
        self.C = cs.clamp_min(1e-6).reciprocal_().mul_(cs > 1e-6)
        self._ax = torch.empty(self.sshape, device=device, dtype=dtype)
        self._bx = torch.empty(self.vshape, device=device, dtype=dtype)

    def iterate(self, x, b, n_iter=1, nonneg=False):
        """Run n_iter SIRT updates on x (in place) for data b; returns x."""
        for _ in range(n_iter):
            self._ax.zero_()
            ops.project(x, self.angles, self.PX, self.z, out=self._ax, du=self.du, dv=self.dv)
            torch.sub(b, self._ax, out=self._ax)                         # b - A x
            self._ax.mul_(self.R)
            self._bx.zero_()
            ops.backproject(self._ax, self.angles, self.SY, self.SX, self.z, out=self._bx,
                            du=self.du, dv=self.dv)
            x.addcmul_(self.C, self._bx)
            if nonneg:
                x.clamp_(min=0)
        return x

    def residual(self, x, b):
        self._ax.zero_()
        ops.project(x, self.angles, self.PX, self.z, out=self._ax, du=self.du, dv=self.dv)
        return float(torch.linalg.vector_norm((self._ax - b).double()))


def gradient_descent(x, b, angles, n_iter, step, PX=None, z_order=0):
    """Minimise 1/2 ||A x - b||^2 with fixed-step gradient descent; returns losses."""
    from .autograd import least_squares_gradient

    losses = []
    for _ in range(n_iter):
        loss, g = least_squares_gradient(x, b, angles, PX, z_order)
        losses.append(float(loss))
        x.sub_(step * g)
    return losses


def lipschitz_step(angles, SY, SX, H, PX=None, z_order=0, n_power=20, device="cuda", seed=0):
    """1 / ||A||^2 by power iteration on AT A (a safe gradient-descent step)."""
    g = torch.Generator(device=device).manual_seed(seed)
    x = torch.randn(ops.vol_shape(SY, SX, H, z_order), device=device, generator=g)
    lam = 1.0
    for _ in range(n_power):
        x /= torch.linalg.vector_norm(x)
        y = ops.project(x, angles, PX, z_order)
        x = ops.backproject(y, angles, SY, SX, z_order)
        lam = float(torch.linalg.vector_norm(x))
    return 1.0 / lam
