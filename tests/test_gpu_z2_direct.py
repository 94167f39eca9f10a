"""The backprojection's z_order 2 volumes written by the 4-slice kernel itself.

For more than 128 slices, vp_backward in z_order 2 transposes only the
sinogram; the kernel adds its sums straight into the (H, SY, SX) volume
(16-byte runs of a tile row when SX is a multiple of 4 and the row is inside
the volume, single floats otherwise).  Its results must equal the z_order 0
operator's, transposed, bit for bit -- with accumulation into a nonzero
volume, ragged tiles, H not a multiple of 256, H slabs of a small caller
workspace and the dv path -- and the FP64 oracle within the float32
tolerance.
"""
import ctypes

import numpy as np
import pytest
import torch

import oracle as vpo
from voxelproj_amd import _lib, ops

pytestmark = pytest.mark.gpu
DEV = "cuda"

# (SY, SX, H, n_angles, PX): SX a multiple of 4 with whole and partial tile
# rows, SX odd (every row on the single-float path), H just above 128 and
# not a multiple of 4 or 256
SHAPES = [(48, 64, 260, 37, 70), (41, 52, 300, 29, 60), (33, 37, 131, 23, 45), (20, 24, 513, 11, 30),
          # degenerate extents: one voxel, one row or column, a detector of one pixel
          (1, 1, 129, 7, 3), (1, 8, 200, 5, 9), (9, 1, 257, 6, 1), (2, 4, 300, 40, 2)]


def _z0_reference(y2, ang, SY, SX, x2_init):
    """AT in z_order 0 on the transposed sinogram, added to the initial
    volume in z_order 0 layout, returned in z_order 2 layout."""
    # (copies: for SY = SX = 1 the permuted volume is already contiguous, and
    # .contiguous() would hand back a view of x2_init)
    y0 = y2.permute(1, 2, 0).clone(memory_format=torch.contiguous_format)       # (A, PX, H)
    x0 = x2_init.permute(1, 2, 0).clone(memory_format=torch.contiguous_format)  # (SY, SX, H)
    x0 = ops.backproject(y0, ang, SY, SX, 0, out=x0)
    return x0.permute(2, 0, 1).contiguous()


@pytest.mark.parametrize("shape", SHAPES)
def test_direct_z2_equals_transposed_z0(shape):
    SY, SX, H, A, PX = shape
    print(f"direct z2 {shape}", flush=True)
    g = torch.Generator(device=DEV).manual_seed(SY * 1000 + H)
    rng = np.random.default_rng(H)
    ang = torch.tensor(rng.uniform(0, np.pi, A).astype(np.float32), device=DEV)
    y2 = torch.randn((H, A, PX), device=DEV, generator=g)
    x_init = torch.randn((H, SY, SX), device=DEV, generator=g)
    ref = _z0_reference(y2, ang, SY, SX, x_init)
    got = ops.backproject(y2, ang, SY, SX, 2, out=x_init.clone())
    assert torch.equal(got, ref)
    # against the FP64 oracle on the sums alone
    sums = (got - x_init).cpu().numpy().astype(np.float64)
    want = vpo.backward_f64(y2.cpu().numpy().astype(np.float64), ang.cpu().numpy(), SY, SX, 2)
    # the sums are taken back out of a volume of N(0,1) values: allow that rounding
    scale = max(np.linalg.norm(want), 1e-30)
    assert np.linalg.norm(sums - want) / scale <= 1e-6 + 2e-7 * np.linalg.norm(x_init.cpu().numpy()) / scale


def test_direct_z2_in_slabs_of_a_caller_workspace():
    SY, SX, H, A, PX = 40, 44, 600, 31, 50
    print("direct z2 slabs", flush=True)
    g = torch.Generator(device=DEV).manual_seed(7)
    ang = torch.linspace(0, 3.1, A, device=DEV)
    y2 = torch.randn((H, A, PX), device=DEV, generator=g)
    x_init = torch.randn((H, SY, SX), device=DEV, generator=g)
    ref = _z0_reference(y2, ang, SY, SX, x_init)
    lib = _lib.load()
    per_slice = SY * SX + A * PX
    for n_ws in (per_slice, 300 * per_slice):  # slabs of 2 slices; of 512 + 88 slices
        ws = torch.empty(n_ws, device=DEV)
        out = x_init.clone()
        rc = lib.vp_backward_ws(2, y2.data_ptr(), out.data_ptr(), ang.data_ptr(), A, SY, SX, H, PX,
                                None, None, 0.0, ws.data_ptr(), n_ws,
                                ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
        assert rc == 0, lib.vp_last_error()
        assert torch.equal(out, ref)


def test_direct_z2_with_dv():
    SY, SX, H, A, PX = 36, 40, 200, 27, 48
    print("direct z2 dv", flush=True)
    rng = np.random.default_rng(11)
    ang = rng.uniform(0, np.pi, A).astype(np.float32)
    du = rng.uniform(-3, 3, A).astype(np.float32)
    dv = rng.uniform(-2.5, 2.5, A).astype(np.float32)
    y2 = rng.standard_normal((H, A, PX)).astype(np.float32)
    got2 = ops.backproject(torch.tensor(y2, device=DEV), ang, SY, SX, 2, du=du, dv=dv).cpu()
    got0 = ops.backproject(torch.tensor(np.ascontiguousarray(y2.transpose(1, 2, 0)), device=DEV), ang,
                           SY, SX, 0, du=du, dv=dv).cpu()
    assert torch.equal(got2, got0.permute(2, 0, 1))
    want = vpo.backward_geom_f64(y2.astype(np.float64), ang, SY, SX, 2, du, dv)
    got = got2.numpy().astype(np.float64)
    assert np.linalg.norm(got - want) / np.linalg.norm(want) <= 1e-6
