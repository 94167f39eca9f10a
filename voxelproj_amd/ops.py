"""Device-tensor operator layer used by autograd, SIRT, batching and serving.

``project(x, angles, PX)`` returns A x and ``backproject(y, angles, SY, SX)``
returns AT y for device tensors, in float32 (hand-written gfx950 kernels) or
float64 (the double-precision kernels, for gradient checks).  Passing ``out``
accumulates into it (out += op(in)), as the reference kernels do; otherwise a
zero-initialised output is allocated.  Everything stays on the device and is
stream-ordered on the current stream.

Layouts (voxelproj/csrc/voxelproj.cpp:15-33):
  z_order 0: x (SY, SX, H), y (A, PX, H)
  z_order 2: x (H, SY, SX), y (H, A, PX)
"""
import torch

from . import _lib

_ENTRY = {
    (torch.float32, "A"): "vp_forward",
    (torch.float32, "AT"): "vp_backward",
    (torch.float64, "A"): "vp_forward_f64",
    (torch.float64, "AT"): "vp_backward_f64",
}


def vol_dims(shape, z_order):
    if z_order == 0:
        SY, SX, H = shape
    elif z_order == 2:
        H, SY, SX = shape
    else:
        raise ValueError(f"z_order must be 0 or 2, got {z_order}")
    return int(SY), int(SX), int(H)


def sino_dims(shape, z_order):
    if z_order == 0:
        A, PX, H = shape
    elif z_order == 2:
        H, A, PX = shape
    else:
        raise ValueError(f"z_order must be 0 or 2, got {z_order}")
    return int(A), int(PX), int(H)


def vol_shape(SY, SX, H, z_order):
    return (SY, SX, H) if z_order == 0 else (H, SY, SX)


def sino_shape(A, PX, H, z_order):
    return (A, PX, H) if z_order == 0 else (H, A, PX)


def device_angles(angles, device):
    """float32 contiguous 1-D angle tensor on `device` (no copy if already so)."""
    if isinstance(angles, torch.Tensor):
        a = angles.to(device=device, dtype=torch.float32)
    else:
        a = torch.as_tensor(angles, dtype=torch.float32, device=device)
    return a.reshape(-1).contiguous()


def _check(t, name):
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise RuntimeError(f"{name} must be a device tensor")
    if t.dtype not in (torch.float32, torch.float64):
        raise RuntimeError(f"{name} must be float32 or float64")
    if not t.is_contiguous():
        raise RuntimeError(f"{name} must be contiguous")


def _launch(kind, z_order, inp, out, angles, A, SY, SX, H, PX):
    lib = _lib.load()
    name = _ENTRY[(inp.dtype, kind)]
    with torch.cuda.device(inp.device):
        stream = torch.cuda.current_stream(inp.device).cuda_stream
        rc = getattr(lib, name)(z_order, inp.data_ptr(), out.data_ptr(), angles.data_ptr(), A,
                                SY, SX, H, PX, stream)
    _lib.check(rc, name)


def project(x, angles, PX=None, z_order=0, out=None):
    """y = A x (or out += A x).  x: device tensor, float32 or float64."""
    _check(x, "x")
    SY, SX, H = vol_dims(x.shape, z_order)
    a = device_angles(angles, x.device)
    PX = max(SY, SX) if PX is None else int(PX)
    shape = sino_shape(a.numel(), PX, H, z_order)
    if out is None:
        out = torch.zeros(shape, dtype=x.dtype, device=x.device)
    else:
        _check(out, "out")
        if tuple(out.shape) != shape or out.dtype != x.dtype:
            raise RuntimeError(f"out must be {x.dtype} of shape {shape}")
    _launch("A", z_order, x, out, a, a.numel(), SY, SX, H, PX)
    return out


def backproject(y, angles, SY=None, SX=None, z_order=0, out=None):
    """x = AT y (or out += AT y).  y: device tensor, float32 or float64."""
    _check(y, "y")
    A, PX, H = sino_dims(y.shape, z_order)
    a = device_angles(angles, y.device)
    if a.numel() != A:
        raise RuntimeError(f"{a.numel()} angles for a sinogram with {A} angle rows")
    SY = PX if SY is None else int(SY)
    SX = PX if SX is None else int(SX)
    shape = vol_shape(SY, SX, H, z_order)
    if out is None:
        out = torch.zeros(shape, dtype=y.dtype, device=y.device)
    else:
        _check(out, "out")
        if tuple(out.shape) != shape or out.dtype != y.dtype:
            raise RuntimeError(f"out must be {y.dtype} of shape {shape}")
    _launch("AT", z_order, y, out, a, A, SY, SX, H, PX)
    return out
