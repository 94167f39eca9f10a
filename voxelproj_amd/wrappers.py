"""Python API: ``forward`` / ``backward`` with the reference's container semantics.

Mirrors voxelproj/wrappers.py:10-122:

* inputs may be numpy arrays, CPU tensors or device tensors; they are
  converted to float32 device tensors (no copy for a float32 device tensor);
* ``y`` / ``x`` (the output argument) may be ``None`` (default shape), a
  shape tuple, or a container that is ACCUMULATED into (``out += op(in)``)
  and returned as the same object;
* with no output container the result comes back in the kind of the input:
  numpy for numpy input, otherwise a tensor on the input's device;
* default shapes: forward z0 ``(len(angles), max(SY,SX), H)``, z2
  ``(H, len(angles), max(SY,SX))``; backward z0 ``(PX, PX, H)``, z2
  ``(H, PX, PX)`` (wrappers.py:40-43, 88-91);
* a container that is not None / tuple / ndarray / Tensor raises TypeError
  (wrappers.py:24); an unknown ``z_order`` raises KeyError (as the
  reference's dict lookup does; the exception here is also a ValueError).

Documented deviations (all turn silent misbehaviour of the reference into an
error or into the intended result):

* shapes are validated before any device work: the input must be 3-D, the
  output shape must match the input (H extent, angle count) -> ValueError;
* ``len(angles)`` must equal the sinogram's angle extent -> ValueError
  (the reference reads past the end of a short angle array);
* a non-contiguous device output is computed in a contiguous buffer and
  copied back (the reference writes it as if it were contiguous);
* any number of angles works (the reference's AT is limited to 512).

``block_size`` is accepted and ignored (no meaning on CDNA).

Extension (keyword-only, absent from the reference): ``du`` / ``dv`` detector
translations along the pixel axis / along H, each None, a number or one value
per angle (see ``voxelproj_amd.ops``).
"""
import numpy as np
import torch

from . import _CU, ops


class ZOrderError(KeyError, ValueError):
    """z_order is not 0 or 2."""


_FORWARD = {0: _CU.forward_z0, 2: _CU.forward_z2}
_BACKWARD = {0: _CU.backward_z0, 2: _CU.backward_z2}


def _device():
    return torch.device("cuda", torch.cuda.current_device())


def gpu_array_with_default_shape(x=None, shape=None):
    """None/tuple -> zeros on the device; ndarray/Tensor -> float32 device tensor."""
    if x is None:
        return torch.zeros(shape, dtype=torch.float32, device="cuda")
    if isinstance(x, tuple):
        return torch.zeros(x, dtype=torch.float32, device="cuda")
    if isinstance(x, np.ndarray):
        return torch.tensor(x, dtype=torch.float32, device="cuda")
    if isinstance(x, torch.Tensor):
        return x.to(dtype=torch.float32, device="cuda")
    raise TypeError("x must be None, a shape tuple, or an array/tensor")


def _shape_of(obj):
    if isinstance(obj, tuple):
        return tuple(int(v) for v in obj)
    return tuple(obj.shape)


def _check_container(obj, name):
    if obj is None or isinstance(obj, (tuple, np.ndarray, torch.Tensor)):
        return
    raise TypeError(f"{name} must be None, a shape tuple, or an array/tensor")


def _check_zorder(z_order):
    if z_order not in (0, 2):
        raise ZOrderError(z_order)


def _n_angles(angles):
    return len(angles)


def forward_shapes(x_shape, n_angles, z_order):
    """Default sinogram shape for a volume (wrappers.py:40-43)."""
    _check_zorder(z_order)
    if len(x_shape) != 3:
        raise ValueError(f"volume must be 3-D, got shape {tuple(x_shape)}")
    if z_order == 0:
        return (n_angles, max(x_shape[0], x_shape[1]), x_shape[2])
    return (x_shape[0], n_angles, max(x_shape[1], x_shape[2]))


def backward_shapes(y_shape, z_order):
    """Default volume shape for a sinogram (wrappers.py:88-91)."""
    _check_zorder(z_order)
    if len(y_shape) != 3:
        raise ValueError(f"sinogram must be 3-D, got shape {tuple(y_shape)}")
    if z_order == 0:
        return (y_shape[1], y_shape[1], y_shape[2])
    return (y_shape[0], y_shape[2], y_shape[2])


def check_pair(x_shape, y_shape, n_angles, z_order):
    """Validate that a volume / sinogram pair is consistent."""
    _check_zorder(z_order)
    if len(x_shape) != 3 or len(y_shape) != 3:
        raise ValueError("volume and sinogram must be 3-D")
    if z_order == 0:
        hx, hy, ay = x_shape[2], y_shape[2], y_shape[0]
    else:
        hx, hy, ay = x_shape[0], y_shape[0], y_shape[1]
    if hx != hy:
        raise ValueError(f"H mismatch: volume {tuple(x_shape)} vs sinogram {tuple(y_shape)}")
    if ay != n_angles:
        raise ValueError(
            f"sinogram {tuple(y_shape)} has {ay} angle rows but {n_angles} angles were given"
        )
    if min(x_shape) < 0 or min(y_shape) < 0:
        raise ValueError("negative extent")


def _device_out(out_obj, shape):
    """Return (device float32 contiguous tensor to accumulate into, needs_copy_back)."""
    if out_obj is None or isinstance(out_obj, tuple):
        return torch.zeros(shape, dtype=torch.float32, device="cuda"), False
    t = gpu_array_with_default_shape(out_obj)
    if not t.is_contiguous():
        t = t.contiguous()
    return t, True


def _finish(out_obj, out_dev, in_obj, is_numpy, device):
    if isinstance(out_obj, (np.ndarray, torch.Tensor)):
        if isinstance(out_obj, np.ndarray):
            np.copyto(out_obj, out_dev.cpu().numpy())
        elif out_obj.data_ptr() != out_dev.data_ptr():
            out_obj.copy_(out_dev.to(out_obj.device))
        return out_obj
    if is_numpy:
        return out_dev.cpu().numpy()
    return out_dev.to(device) if device is not None else out_dev


def _sino_px(y_shape, z_order):
    return y_shape[1] if z_order == 0 else y_shape[2]


def _angles_dev(angles):
    if isinstance(angles, torch.Tensor):
        a = angles.to(dtype=torch.float32, device="cuda")
    else:
        a = torch.tensor(np.asarray(angles, dtype=np.float32), device="cuda")
    return a.reshape(-1).contiguous()


def forward(x, angles, y, z_order=0, block_size=[8, 8, 8], *, du=None, dv=None):  # noqa: B006
    """Forward projection y (+)= A x.  See the module docstring for semantics."""
    _check_zorder(z_order)
    _check_container(y, "y")
    if not isinstance(x, (np.ndarray, torch.Tensor)):
        raise TypeError("x must be an array/tensor")
    is_numpy = isinstance(x, np.ndarray)
    device = x.device if isinstance(x, torch.Tensor) else None
    n = _n_angles(angles)
    shape = forward_shapes(tuple(x.shape), n, z_order)
    y_shape = shape if y is None else _shape_of(y)
    check_pair(tuple(x.shape), y_shape, n, z_order)

    x_dev = gpu_array_with_default_shape(x)
    a_dev = _angles_dev(angles)
    y_dev, _ = _device_out(y, y_shape)
    if du is None and dv is None:
        _FORWARD[z_order](x_dev, y_dev, a_dev, block_size=block_size)
    else:
        ops.project(x_dev.contiguous(), a_dev, _sino_px(y_shape, z_order), z_order, out=y_dev,
                    du=du, dv=dv)
    return _finish(y, y_dev, x, is_numpy, device)


def backward(y, angles, x, z_order=0, block_size=[8, 8, 8], *, du=None, dv=None):  # noqa: B006
    """Backprojection x (+)= AT y.  See the module docstring for semantics."""
    _check_zorder(z_order)
    _check_container(x, "x")
    if not isinstance(y, (np.ndarray, torch.Tensor)):
        raise TypeError("y must be an array/tensor")
    is_numpy = isinstance(y, np.ndarray)
    device = y.device if isinstance(y, torch.Tensor) else None
    n = _n_angles(angles)
    shape = backward_shapes(tuple(y.shape), z_order)
    x_shape = shape if x is None else _shape_of(x)
    check_pair(x_shape, tuple(y.shape), n, z_order)

    y_dev = gpu_array_with_default_shape(y)
    a_dev = _angles_dev(angles)
    x_dev, _ = _device_out(x, x_shape)
    if du is None and dv is None:
        _BACKWARD[z_order](y_dev, x_dev, a_dev, block_size=block_size)
    else:
        SY, SX, _ = ops.vol_dims(x_shape, z_order)
        ops.backproject(y_dev.contiguous(), a_dev, SY, SX, z_order, out=x_dev, du=du, dv=dv)
    return _finish(x, x_dev, y, is_numpy, device)
