"""Drop-in for the reference's compiled module ``voxelproj._CU``.

Same four entry points, argument order and defaults as the pybind module
(voxelproj/csrc/voxelproj.cpp:11-34):

    forward_z0(in, out, angles, block_size=[8, 8, 8])   # out += A in,  H fastest
    backward_z0(in, out, angles, block_size=[8, 8, 8])  # out += AT in, H fastest
    forward_z2(in, out, angles, block_size=[8, 8, 8])   # H slowest
    backward_z2(in, out, angles, block_size=[8, 8, 8])

All tensors are device (HIP) tensors.  ``in`` is checked exactly like the
reference's TORCH_CHECKs (projectors.cu:36-40): RuntimeError "Tensor must be
on CUDA", "Only float32 tensors supported", "Tensor must be contiguous",
"Tensor must be 3D".  Unlike the reference, ``out`` and ``angles`` are
validated too (device, float32, contiguous, matching extents, one angle per
sinogram row) instead of being read or written out of bounds.

``block_size`` is accepted for compatibility; CUDA thread-block shapes have
no meaning for these kernels, so it is ignored.
"""
import torch

from . import _lib


def _check_in(t):
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise RuntimeError("Tensor must be on CUDA")
    if t.dtype != torch.float32:
        raise RuntimeError("Only float32 tensors supported")
    if not t.is_contiguous():
        raise RuntimeError("Tensor must be contiguous")
    if t.dim() != 3:
        raise RuntimeError("Tensor must be 3D")


def _check_out(t, shape, like):
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise RuntimeError("Output tensor must be on CUDA")
    if t.dtype != torch.float32:
        raise RuntimeError("Output tensor must be float32")
    if not t.is_contiguous():
        raise RuntimeError("Output tensor must be contiguous")
    if t.dim() != 3:
        raise RuntimeError("Output tensor must be 3D")
    if t.device != like.device:
        raise RuntimeError("Input and output tensors must be on the same device")


def _check_angles(a, n, like):
    if not isinstance(a, torch.Tensor) or not a.is_cuda or a.dtype != torch.float32:
        raise RuntimeError("angles must be a float32 CUDA tensor")
    if not a.is_contiguous():
        raise RuntimeError("angles must be contiguous")
    if a.numel() != n:
        raise RuntimeError(
            f"angles has {a.numel()} entries but the sinogram has {n} angle rows"
        )
    if a.device != like.device:
        raise RuntimeError("angles must be on the same device as the data")


def _stream(t):
    return ctypes_stream(torch.cuda.current_stream(t.device))


def ctypes_stream(stream):
    return stream.cuda_stream


def _call(name, z, inp, out, angles, A, SY, SX, H, PX):
    lib = _lib.load()
    with torch.cuda.device(inp.device):
        stream = _stream(inp)
        if z == 0:
            rc = getattr(lib, name)(z, inp.data_ptr(), out.data_ptr(), angles.data_ptr(), A,
                                    SY, SX, H, PX, stream)
        else:
            # z_order 2 runs through layout transposes; their workspace comes
            # from torch's caching allocator (stream-ordered, reused across calls)
            ws = _lib.workspace(0, z, A, SY, SX, H, PX, False, inp.device)
            fn = lib.vp_forward_ws if name == "vp_forward" else lib.vp_backward_ws
            rc = fn(z, inp.data_ptr(), out.data_ptr(), angles.data_ptr(), A, SY, SX, H, PX,
                    None, None, 0.0, ws.data_ptr(), ws.numel(), stream)
    _lib.check(rc, name)


def forward_z0(in_, out, angles, block_size=[8, 8, 8]):  # noqa: B006 - mirrors the reference
    """out[a, p, h] += sum_v W_a[p, v] in[v, h]; in (SY, SX, H), out (A, PX, H)."""
    _check_in(in_)
    SY, SX, H = in_.shape
    _check_out(out, None, in_)
    A, PX, Ho = out.shape
    if Ho != H:
        raise RuntimeError(f"H mismatch: volume has {H} slices, sinogram {Ho}")
    _check_angles(angles, A, in_)
    _call("vp_forward", 0, in_, out, angles, A, SY, SX, H, PX)


def forward_z2(in_, out, angles, block_size=[8, 8, 8]):  # noqa: B006
    """in (H, SY, SX), out (H, A, PX)."""
    _check_in(in_)
    H, SY, SX = in_.shape
    _check_out(out, None, in_)
    Ho, A, PX = out.shape
    if Ho != H:
        raise RuntimeError(f"H mismatch: volume has {H} slices, sinogram {Ho}")
    _check_angles(angles, A, in_)
    _call("vp_forward", 2, in_, out, angles, A, SY, SX, H, PX)


def backward_z0(in_, out, angles, block_size=[8, 8, 8]):  # noqa: B006
    """out[v, h] += sum_a sum_p W_a[p, v] in[a, p, h]; in (A, PX, H), out (SY, SX, H)."""
    _check_in(in_)
    A, PX, H = in_.shape
    _check_out(out, None, in_)
    SY, SX, Ho = out.shape
    if Ho != H:
        raise RuntimeError(f"H mismatch: sinogram has {H} slices, volume {Ho}")
    _check_angles(angles, A, in_)
    _call("vp_backward", 0, in_, out, angles, A, SY, SX, H, PX)


def backward_z2(in_, out, angles, block_size=[8, 8, 8]):  # noqa: B006
    """in (H, A, PX), out (H, SY, SX)."""
    _check_in(in_)
    H, A, PX = in_.shape
    _check_out(out, None, in_)
    Ho, SY, SX = out.shape
    if Ho != H:
        raise RuntimeError(f"H mismatch: sinogram has {H} slices, volume {Ho}")
    _check_angles(angles, A, in_)
    _call("vp_backward", 2, in_, out, angles, A, SY, SX, H, PX)
