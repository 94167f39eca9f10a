"""GPU parity of the translated detector geometry (du along the detector, dv
along H) and of the H-slab streaming of host-resident data (chunked.py).

The translations are an extension (the reference lists them as future work,
README.md:65-68); the checker is the FP64 geometry oracle
(oracle.forward_geom_f64 / backward_geom_f64), itself pinned in
tests/test_geometry_oracle.py against closed forms and the polygon brute
force.  Tolerances as for the plain operator: rel-L2 <= 1e-6,
max|err| / max|ref| <= 1e-5, adjoint gap <= 1e-6 (float32); float64 paths
within 1e-12.
"""
import numpy as np
import pytest
import torch

import oracle as vpo
import voxelproj_amd as vp
from voxelproj_amd import autograd, chunked, ops

pytestmark = pytest.mark.gpu

DEV = "cuda"
REL_L2 = 1e-6
MAX_REL = 1e-5


def _err(got, ref):
    got = np.asarray(got, np.float64)
    return (float(np.linalg.norm(got - ref) / np.linalg.norm(ref)),
            float(np.abs(got - ref).max() / np.abs(ref).max()))


def _close(got, ref, what):
    r, m = _err(got, ref)
    print(f"{what}: rel-L2 {r:.2e} max-rel {m:.2e}", flush=True)
    assert r <= REL_L2 and m <= MAX_REL, what


CASES = [
    # (SY, SX, H, A, PX, du, dv)
    (24, 20, 9, 30, 32, 0.37, None),
    (24, 20, 9, 30, 32, None, -1.6),
    (31, 33, 13, 45, 40, "per-angle", "per-angle"),
    (64, 64, 70, 100, 80, "per-angle", 2.25),
    (5, 6, 3, 7, 11, -7.5, 0.5),
]


def _shifts(case, A, rng):
    du, dv = case[5], case[6]
    if du == "per-angle":
        du = rng.uniform(-6, 6, A).astype(np.float32)
    if dv == "per-angle":
        dv = rng.uniform(-3, 3, A).astype(np.float32)
    return du, dv


@pytest.mark.parametrize("z_order", [0, 2])
@pytest.mark.parametrize("case", CASES)
def test_translated_pair_matches_oracle(case, z_order):
    rng = np.random.default_rng(11)
    SY, SX, H, A, PX = case[:5]
    du, dv = _shifts(case, A, rng)
    ang = np.linspace(0, np.pi, A, endpoint=False).astype(np.float32)
    x = rng.standard_normal(vpo.x_shape(z_order, SY, SX, H)).astype(np.float32)
    y = rng.standard_normal(vpo.y_shape(z_order, A, PX, H)).astype(np.float32)
    print(f"case {case[:5]} z{z_order}", flush=True)
    ax = ops.project(torch.tensor(x, device=DEV), ang, PX, z_order, du=du, dv=dv).cpu().numpy()
    aty = ops.backproject(torch.tensor(y, device=DEV), ang, SY, SX, z_order, du=du, dv=dv).cpu().numpy()
    _close(ax, vpo.forward_geom_f64(x, ang, PX, z_order, du, dv), "A")
    _close(aty, vpo.backward_geom_f64(y, ang, SY, SX, z_order, du, dv), "AT")
    gap = abs(np.vdot(ax.astype(np.float64), y) - np.vdot(x.astype(np.float64), aty))
    gap /= np.linalg.norm(ax) * np.linalg.norm(y)
    assert gap <= 1e-6


@pytest.mark.parametrize("z_order", [0, 2])
def test_translated_f64_matches_oracle(z_order):
    rng = np.random.default_rng(3)
    SY, SX, H, A, PX = 14, 17, 6, 19, 21
    ang = np.linspace(0, np.pi, A, endpoint=False).astype(np.float32)
    du = rng.uniform(-3, 3, A).astype(np.float32)
    dv = rng.uniform(-2, 2, A).astype(np.float32)
    x = rng.standard_normal(vpo.x_shape(z_order, SY, SX, H))
    y = rng.standard_normal(vpo.y_shape(z_order, A, PX, H))
    ax = ops.project(torch.tensor(x, device=DEV), ang, PX, z_order, du=du, dv=dv).cpu().numpy()
    aty = ops.backproject(torch.tensor(y, device=DEV), ang, SY, SX, z_order, du=du, dv=dv).cpu().numpy()
    assert np.allclose(ax, vpo.forward_geom_f64(x, ang, PX, z_order, du, dv), rtol=0, atol=1e-12)
    assert np.allclose(aty, vpo.backward_geom_f64(y, ang, SY, SX, z_order, du, dv), rtol=0, atol=1e-12)


@pytest.mark.parametrize("z_order", [0, 2])
def test_zero_translation_is_bitwise_plain(z_order):
    rng = np.random.default_rng(5)
    SY, SX, H, A, PX = 40, 36, 70, 50, 48
    ang = np.linspace(0, np.pi, A).astype(np.float32)
    x = torch.tensor(rng.standard_normal(vpo.x_shape(z_order, SY, SX, H)).astype(np.float32), device=DEV)
    y = torch.tensor(rng.standard_normal(vpo.y_shape(z_order, A, PX, H)).astype(np.float32), device=DEV)
    z = np.zeros(A, np.float32)
    assert torch.equal(ops.project(x, ang, PX, z_order), ops.project(x, ang, PX, z_order, du=z, dv=z))
    assert torch.equal(ops.backproject(y, ang, SY, SX, z_order),
                       ops.backproject(y, ang, SY, SX, z_order, du=z, dv=z))


def test_wrappers_accept_translations():
    rng = np.random.default_rng(9)
    SY, SX, H, A = 16, 16, 5, 12
    ang = np.linspace(0, np.pi, A, endpoint=False).astype(np.float32)
    x = rng.standard_normal((SY, SX, H)).astype(np.float32)
    y = vp.forward(x, ang, None, du=0.75, dv=-0.5)
    assert isinstance(y, np.ndarray)
    _close(y, vpo.forward_geom_f64(x, ang, 16, 0, 0.75, -0.5), "wrappers.forward")
    xb = vp.backward(y, ang, (SY, SX, H), du=0.75, dv=-0.5)
    _close(xb.cpu().numpy() if isinstance(xb, torch.Tensor) else xb,
           vpo.backward_geom_f64(y, ang, SY, SX, 0, 0.75, -0.5), "wrappers.backward")


def test_volume_offset_helper():
    rng = np.random.default_rng(2)
    SY = SX = 20
    H, A, PX = 4, 24, 30
    ang = np.linspace(0, np.pi, A, endpoint=False).astype(np.float32)
    x = np.zeros((SY, SX, H), np.float32)
    x[5:12, 6:14] = rng.standard_normal((7, 8, H))
    xs = np.zeros_like(x)
    xs[5 + 3:12 + 3, 6 - 2:14 - 2] = x[5:12, 6:14]  # moved by dx = -2, dy = +3
    du = ops.volume_offset_du(ang, dx=-2, dy=3)
    got = ops.project(torch.tensor(x, device=DEV), ang, PX, du=du).cpu().numpy()
    ref = ops.project(torch.tensor(xs, device=DEV), ang, PX).cpu().numpy()
    assert np.abs(got - ref).max() <= 1e-5 * np.abs(ref).max()


def test_gradcheck_with_translations():
    torch.manual_seed(0)
    SY, SX, H, A, PX = 8, 7, 3, 6, 11
    ang = torch.tensor(np.linspace(0, np.pi, A, endpoint=False).astype(np.float32), device=DEV)
    du = np.linspace(-1.3, 2.1, A).astype(np.float32)
    x = torch.randn((SY, SX, H), dtype=torch.float64, device=DEV, requires_grad=True)
    assert torch.autograd.gradcheck(lambda t: autograd.project(t, ang, PX, 0, du, 0.4), (x,),
                                    eps=1e-6, atol=1e-6)
    y = torch.randn((A, PX, H), dtype=torch.float64, device=DEV, requires_grad=True)
    assert torch.autograd.gradcheck(lambda t: autograd.backproject(t, ang, SY, SX, 0, du, -1.7), (y,),
                                    eps=1e-6, atol=1e-6)


@pytest.mark.parametrize("z_order", [0, 2])
@pytest.mark.parametrize("dv", [None, 1.5, "per-angle"])
def test_chunked_equals_in_memory(z_order, dv):
    rng = np.random.default_rng(4)
    SY, SX, H, A, PX = 48, 40, 37, 33, 52
    ang = np.linspace(0, np.pi, A, endpoint=False).astype(np.float32)
    if dv == "per-angle":
        dv = rng.uniform(-2.5, 2.5, A).astype(np.float32)
    du = rng.uniform(-2, 2, A).astype(np.float32)
    x = rng.standard_normal(vpo.x_shape(z_order, SY, SX, H)).astype(np.float32)
    y = rng.standard_normal(vpo.y_shape(z_order, A, PX, H)).astype(np.float32)
    slice_bytes = 4 * (SY * SX + 2 * A * PX)
    budget = 2 * 3 * slice_bytes * 12  # about 6-10 slices per slab: several slabs with halos
    print(f"chunked z{z_order} dv={'arr' if isinstance(dv, np.ndarray) else dv}", flush=True)
    ref_y = ops.project(torch.tensor(x, device=DEV), ang, PX, z_order, du=du, dv=dv).cpu().numpy()
    got_y = chunked.forward(x, ang, None, PX, z_order, du=du, dv=dv, max_device_bytes=budget)
    assert np.allclose(got_y, ref_y, rtol=1e-6, atol=1e-5 * np.abs(ref_y).max())
    ref_x = ops.backproject(torch.tensor(y, device=DEV), ang, SY, SX, z_order, du=du, dv=dv).cpu().numpy()
    acc = np.ones(vpo.x_shape(z_order, SY, SX, H), np.float32)
    got_x = chunked.backward(y, ang, acc, SY, SX, z_order, du=du, dv=dv, max_device_bytes=budget)
    assert got_x is acc  # accumulates into the given host array
    assert np.allclose(got_x - 1.0, ref_x, rtol=1e-6, atol=1e-5 * np.abs(ref_x).max())


def test_chunked_matches_oracle():
    rng = np.random.default_rng(8)
    SY, SX, H, A, PX = 30, 30, 20, 24, 36
    ang = np.linspace(0, np.pi, A, endpoint=False).astype(np.float32)
    x = rng.standard_normal((SY, SX, H)).astype(np.float32)
    budget = 2 * 3 * 4 * (SY * SX + 2 * A * PX) * 6
    got = chunked.forward(x, ang, None, PX, 0, du=0.8, dv=-2.2, max_device_bytes=budget)
    _close(got, vpo.forward_geom_f64(x, ang, PX, 0, 0.8, -2.2), "chunked A")
