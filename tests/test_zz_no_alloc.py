"""vp_project_noalloc: the allocation-free variant of vp_forward / vp_backward.

Called through the C ABI and compared with the existing entry points on two
calls in a row, the second with a smaller input: z_order 2 (which runs
through the static scratch, in several H slabs for the first call) and
z_order 0.  The bar is the projectors' tolerance against the FP64 operator
(rel-L2 <= 1e-6), applied to the difference between the two entry points;
the first call is also checked against the FP64 oracle on a few slices.
"""
import numpy as np
import pytest
import torch

import oracle as vpo
from voxelproj_amd import _lib

pytestmark = pytest.mark.gpu
DEV = "cuda"


def rel_l2(a, b):
    a = a.double()
    b = b.double()
    return float((a - b).norm() / b.norm())


def _stream():
    return torch.cuda.current_stream().cuda_stream


def _existing(lib, op, z, inp, out, ang, dims):
    A, SY, SX, H, PX = dims
    fn = lib.vp_forward if op == 0 else lib.vp_backward
    _lib.check(fn(z, inp.data_ptr(), out.data_ptr(), ang.data_ptr(), A, SY, SX, H, PX, _stream()),
               "existing entry point")


def _noalloc(lib, op, z, inp, out, ang, dims):
    A, SY, SX, H, PX = dims
    _lib.check(lib.vp_project_noalloc(op, z, inp.data_ptr(), out.data_ptr(), ang.data_ptr(), A,
                                      SY, SX, H, PX, _stream()), "vp_project_noalloc")


def _shapes(z, dims):
    A, SY, SX, H, PX = dims
    vol = (SY, SX, H) if z == 0 else (H, SY, SX)
    sino = (A, PX, H) if z == 0 else (H, A, PX)
    return vol, sino


@pytest.mark.parametrize("z", [2, 0])
def test_noalloc_matches_existing_entry_points(z):
    lib = _lib.load()
    scratch = lib.vp_noalloc_scratch_elems()
    big = (192, 256, 256, 300, 256)     # z_order 2: 114688 floats per slice -> 3 slabs
    small = (45, 64, 64, 40, 64)        # fits one slab
    assert 300 * (256 * 256 + 192 * 256) > scratch > 256 * 256 + 192 * 256
    g = torch.Generator(device=DEV).manual_seed(17)
    for k, dims in enumerate((big, small)):
        A, SY, SX, H, PX = dims
        ang = torch.tensor(np.linspace(0, np.pi, A, endpoint=False).astype(np.float32), device=DEV)
        vol, sino = _shapes(z, dims)
        x = torch.randn(vol, device=DEV, generator=g)
        y = torch.randn(sino, device=DEV, generator=g)
        # accumulate semantics: both entry points add into a prefilled output
        y0 = torch.randn(sino, device=DEV, generator=g)
        x0 = torch.randn(vol, device=DEV, generator=g)
        ya_ref, ya = y0.clone(), y0.clone()
        xb_ref, xb = x0.clone(), x0.clone()
        _existing(lib, 0, z, x, ya_ref, ang, dims)
        _noalloc(lib, 0, z, x, ya, ang, dims)
        _existing(lib, 1, z, y, xb_ref, ang, dims)
        _noalloc(lib, 1, z, y, xb, ang, dims)
        torch.cuda.synchronize()
        ea, eb = rel_l2(ya - y0, ya_ref - y0), rel_l2(xb - x0, xb_ref - x0)
        print(f"call {k} z{z} {dims}: noalloc vs existing rel-L2 A {ea:.2e}, AT {eb:.2e}", flush=True)
        assert ea <= 1e-6 and eb <= 1e-6
        if k == 0:
            an = ang.cpu().numpy()
            xn, yn = x.cpu().numpy(), y.cpu().numpy()
            da, db = (ya - y0).cpu().numpy(), (xb - x0).cpu().numpy()
            for h in (0, H // 2, H - 1):
                ref = vpo.forward_f64(xn, an, PX, z, h, h + 1)
                refb = vpo.backward_f64(yn, an, SY, SX, z, h, h + 1)
                sl = (slice(None), slice(None), h) if z == 0 else (h,)
                assert np.linalg.norm(da[sl] - ref[sl]) <= 1e-6 * np.linalg.norm(ref[sl]), h
                assert np.linalg.norm(db[sl] - refb[sl]) <= 1e-6 * np.linalg.norm(refb[sl]), h


def test_noalloc_rejects_problems_beyond_its_scratch():
    lib = _lib.load()
    n = lib.vp_noalloc_scratch_elems()
    side = int(np.sqrt(n)) + 1           # one slice of the volume alone exceeds the scratch
    p = torch.zeros(4, device=DEV)
    rc = lib.vp_project_noalloc(0, 2, p.data_ptr(), p.data_ptr(), p.data_ptr(), 1, side, side, 1,
                                side, _stream())
    assert rc == -2
    assert b"scratch" in lib.vp_last_error()
