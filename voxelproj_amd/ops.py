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

Detector translations (an extension; the reference lists them as future work,
README.md:65-68): ``du`` shifts the detector along its pixel axis (pixels),
``dv`` along H (slices, linear interpolation between slices, zero outside the
volume).  Each is None, one number, or one value per angle.  The pair stays
exactly adjoint.  ``volume_offset_du`` converts an in-plane volume translation
into the equivalent per-angle ``du``.
"""
import numpy as np
import torch

from . import _lib

_ENTRY = {
    (torch.float32, "A"): "vp_forward",
    (torch.float32, "AT"): "vp_backward",
    (torch.float64, "A"): "vp_forward_f64",
    (torch.float64, "AT"): "vp_backward_f64",
}
_ENTRY_GEOM = {
    (torch.float32, "A"): "vp_forward_geom",
    (torch.float32, "AT"): "vp_backward_geom",
    (torch.float64, "A"): "vp_forward_geom_f64",
    (torch.float64, "AT"): "vp_backward_geom_f64",
}


class Shift:
    """A per-angle detector translation on the device, with max |value|."""

    def __init__(self, values, n, device, name):
        if isinstance(values, torch.Tensor):
            t = values.detach().to(dtype=torch.float32).reshape(-1)
        else:
            t = torch.as_tensor(np.asarray(values, dtype=np.float32).reshape(-1))
        if t.numel() == 1:
            t = t.expand(n)
        if t.numel() != n:
            raise ValueError(f"{name} needs 1 or {n} values, got {t.numel()}")
        host = t.cpu()
        if not bool(torch.isfinite(host).all()):
            raise ValueError(f"{name} must be finite")
        self.max_abs = float(host.abs().max()) if n else 0.0
        if self.max_abs > 1.0e6:
            raise ValueError(f"|{name}| must be <= 1e6")
        self.tensor = t.to(device=device).contiguous()


def make_shifts(du, dv, n, device):
    """(Shift or None, Shift or None) for the given translations."""
    su = None if du is None else (du if isinstance(du, Shift) else Shift(du, n, device, "du"))
    sv = None if dv is None else (dv if isinstance(dv, Shift) else Shift(dv, n, device, "dv"))
    return su, sv


def volume_offset_du(angles, dx=0.0, dy=0.0):
    """Per-angle du equivalent to translating the volume by (dx, dy) voxels
    in its plane: the content moves by dx cos + dy sin on the detector, i.e.
    du = -(dx cos(phi) + dy sin(phi))."""
    a = np.asarray(angles.cpu() if isinstance(angles, torch.Tensor) else angles, dtype=np.float64)
    return -(dx * np.cos(a) + dy * np.sin(a))


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


def _launch(kind, z_order, inp, out, angles, A, SY, SX, H, PX, du=None, dv=None):
    lib = _lib.load()
    with torch.cuda.device(inp.device):
        stream = torch.cuda.current_stream(inp.device).cuda_stream
        du_p = du.tensor.data_ptr() if du is not None else None
        dv_p = dv.tensor.data_ptr() if dv is not None else None
        du_max = du.max_abs if du is not None else 0.0
        if inp.dtype == torch.float32 and (z_order == 2 or dv is not None):
            # float32 paths that need a workspace take it from torch's allocator
            name = "vp_forward_ws" if kind == "A" else "vp_backward_ws"
            ws = _lib.workspace(False, z_order, A, SY, SX, H, PX, dv is not None, inp.device)
            rc = getattr(lib, name)(z_order, inp.data_ptr(), out.data_ptr(), angles.data_ptr(),
                                    A, SY, SX, H, PX, du_p, dv_p, du_max, ws.data_ptr(),
                                    ws.numel(), stream)
        elif du is None and dv is None:
            name = _ENTRY[(inp.dtype, kind)]
            rc = getattr(lib, name)(z_order, inp.data_ptr(), out.data_ptr(), angles.data_ptr(),
                                    A, SY, SX, H, PX, stream)
        else:
            name = _ENTRY_GEOM[(inp.dtype, kind)]
            rc = getattr(lib, name)(z_order, inp.data_ptr(), out.data_ptr(), angles.data_ptr(),
                                    A, SY, SX, H, PX, du_p, dv_p, du_max, stream)
    _lib.check(rc, name)


def project(x, angles, PX=None, z_order=0, out=None, du=None, dv=None):
    """y = A x (or out += A x).  x: device tensor, float32 or float64;
    du / dv: optional detector translations (see the module docstring)."""
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
    su, sv = make_shifts(du, dv, a.numel(), x.device)
    _launch("A", z_order, x, out, a, a.numel(), SY, SX, H, PX, su, sv)
    return out


def backproject(y, angles, SY=None, SX=None, z_order=0, out=None, du=None, dv=None):
    """x = AT y (or out += AT y).  y: device tensor, float32 or float64;
    du / dv: optional detector translations (see the module docstring)."""
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
    su, sv = make_shifts(du, dv, A, y.device)
    _launch("AT", z_order, y, out, a, A, SY, SX, H, PX, su, sv)
    return out
