"""Workspace handling of the z_order 2 (and dv) paths on the GPU.

* a caller's workspace smaller than the whole problem makes z_order 2 run in
  H slabs: results match the whole-workspace call and the FP64 oracle;
* the library's cached workspace (vp_forward / vp_backward called directly)
  gives the same bits as the torch-workspace path, also after it has grown;
* vp_release_workspaces frees the cache and later calls still work.
"""
import ctypes

import numpy as np
import pytest
import torch

import oracle as vpo
from voxelproj_amd import _CU, _lib

pytestmark = pytest.mark.gpu
DEV = "cuda"


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def _stream():
    return torch.cuda.current_stream().cuda_stream


def _ws_call(op, inp, out, ang, dims, n_ws):
    lib = _lib.load()
    A, SY, SX, H, PX = dims
    ws = torch.empty(max(n_ws, 1), device=DEV)
    fn = lib.vp_forward_ws if op == "A" else lib.vp_backward_ws
    rc = fn(2, inp.data_ptr(), out.data_ptr(), ang.data_ptr(), A, SY, SX, H, PX, None, None, 0.0,
            ws.data_ptr(), n_ws, _stream())
    _lib.check(rc, "ws call")
    return out


@pytest.mark.parametrize("slab", ["one slice", "256 slices", "whole"])
def test_z2_in_slabs(slab):
    A, SY, SX, H, PX = 33, 40, 52, 700, 60
    dims = (A, SY, SX, H, PX)
    g = torch.Generator(device=DEV).manual_seed(8)
    ang = torch.tensor(np.linspace(0, np.pi, A, endpoint=False).astype(np.float32), device=DEV)
    x = torch.randn((H, SY, SX), device=DEV, generator=g)
    y = torch.randn((H, A, PX), device=DEV, generator=g)
    per = SY * SX + A * PX
    n_ws = {"one slice": per, "256 slices": 300 * per, "whole": H * per}[slab]
    ya = _ws_call("A", x, torch.zeros((H, A, PX), device=DEV), ang, dims, n_ws)
    xb = _ws_call("AT", y, torch.zeros((H, SY, SX), device=DEV), ang, dims, n_ws)
    ya_ref = torch.zeros_like(ya)
    _CU.forward_z2(x, ya_ref, ang)
    xb_ref = torch.zeros_like(xb)
    _CU.backward_z2(y, xb_ref, ang)
    torch.testing.assert_close(ya, ya_ref, rtol=0, atol=1e-5 * float(ya_ref.abs().max()))
    torch.testing.assert_close(xb, xb_ref, rtol=0, atol=1e-5 * float(xb_ref.abs().max()))
    an = ang.cpu().numpy()
    for h in (0, 255, 256, 699):
        ref = vpo.forward_f64(x[h:h + 1].cpu().numpy(), an, PX, 2)
        assert rel_l2(ya[h:h + 1].cpu().numpy(), ref) <= 1e-6, h
        refb = vpo.backward_f64(y[h:h + 1].cpu().numpy(), an, SY, SX, 2)
        assert rel_l2(xb[h:h + 1].cpu().numpy(), refb) <= 1e-6, h


def test_cached_workspace_grows_and_matches():
    lib = _lib.load()
    g = torch.Generator(device=DEV).manual_seed(9)
    for (A, SY, SX, H, PX) in ((20, 24, 30, 40, 32), (45, 60, 50, 300, 64), (20, 24, 30, 40, 32)):
        ang = torch.tensor(np.linspace(0, np.pi, A, endpoint=False).astype(np.float32), device=DEV)
        x = torch.randn((H, SY, SX), device=DEV, generator=g)
        y = torch.randn((H, A, PX), device=DEV, generator=g)
        got_a = torch.zeros((H, A, PX), device=DEV)
        got_b = torch.zeros((H, SY, SX), device=DEV)
        _lib.check(lib.vp_forward(2, x.data_ptr(), got_a.data_ptr(), ang.data_ptr(), A, SY, SX, H,
                                  PX, _stream()), "vp_forward")
        _lib.check(lib.vp_backward(2, y.data_ptr(), got_b.data_ptr(), ang.data_ptr(), A, SY, SX, H,
                                   PX, _stream()), "vp_backward")
        ref_a = torch.zeros_like(got_a)
        _CU.forward_z2(x, ref_a, ang)
        ref_b = torch.zeros_like(got_b)
        _CU.backward_z2(y, ref_b, ang)
        assert torch.equal(got_a, ref_a)
        assert torch.equal(got_b, ref_b)
    torch.cuda.synchronize()
    assert lib.vp_release_workspaces() == 0
    got_a.zero_()
    _lib.check(lib.vp_forward(2, x.data_ptr(), got_a.data_ptr(), ang.data_ptr(), A, SY, SX, H,
                              PX, _stream()), "vp_forward after release")
    assert torch.equal(got_a, ref_a)
