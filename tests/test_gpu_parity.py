"""GPU parity: the HIP projectors against the FP64 oracle, the golden fixtures
and size-independent properties at the full benchmark sizes.

Tolerances (float32 results vs the exact FP64 operator):
  rel-L2 <= 1e-6, max|err| / max|ref| <= 1e-5, adjoint gap <= 1e-6.
"""
import glob
import os

import numpy as np
import pytest
import torch

import oracle as vpo
import voxelproj_amd as vp
from voxelproj_amd import _CU

pytestmark = pytest.mark.gpu

REL_L2 = 1e-6
MAX_REL = 1e-5
ADJ = 1e-6
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
DEV = "cuda"


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def max_rel(a, b):
    return float(np.abs(np.asarray(a, np.float64) - b).max() / np.abs(b).max())


def assert_close(got, ref, what=""):
    r, m = rel_l2(got, ref), max_rel(got, ref)
    print(f"{what}: rel-L2 {r:.2e} max-rel {m:.2e}", flush=True)
    assert r <= REL_L2, f"{what} rel-L2 {r:.3e}"
    assert m <= MAX_REL, f"{what} max-rel {m:.3e}"


def gpu_forward(x, angles, PX, z):
    xt = torch.as_tensor(np.ascontiguousarray(x), device=DEV)
    at = torch.as_tensor(np.asarray(angles, np.float32), device=DEV)
    if z == 0:
        SY, SX, H = xt.shape
        y = torch.zeros((len(angles), PX, H), device=DEV)
        _CU.forward_z0(xt, y, at)
    else:
        H, SY, SX = xt.shape
        y = torch.zeros((H, len(angles), PX), device=DEV)
        _CU.forward_z2(xt, y, at)
    return y.cpu().numpy()


def gpu_backward(y, angles, SY, SX, z):
    yt = torch.as_tensor(np.ascontiguousarray(y), device=DEV)
    at = torch.as_tensor(np.asarray(angles, np.float32), device=DEV)
    if z == 0:
        A, PX, H = yt.shape
        x = torch.zeros((SY, SX, H), device=DEV)
        _CU.backward_z0(yt, x, at)
    else:
        H, A, PX = yt.shape
        x = torch.zeros((H, SY, SX), device=DEV)
        _CU.backward_z2(yt, x, at)
    return x.cpu().numpy()


def to_z(a, z):
    return a if z == 0 else np.ascontiguousarray(np.transpose(a, (2, 0, 1)))


def from_z(a, z):
    return a if z == 0 else np.ascontiguousarray(np.transpose(a, (1, 2, 0)))


def seeded(shape, seed):
    g = torch.Generator().manual_seed(seed)
    return torch.randn(shape, generator=g).numpy()


# ---------------------------------------------------------------- fixtures
FIXTURES = sorted(p for p in glob.glob(os.path.join(GOLDEN, "*.npz")) if "weight_sums" not in p)


@pytest.mark.parametrize("z", [0, 2])
@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p)[:-4] for p in FIXTURES])
def test_golden_fixture(path, z):
    d = np.load(path)
    SY, SX, H, A, PX = (int(v) for v in d["geom"])
    print(f"fixture {os.path.basename(path)} z{z}", flush=True)
    yg = from_z(gpu_forward(to_z(d["x"], z), d["angles"], PX, z), z)
    assert_close(yg, d["Ax"], "A")
    xg = from_z(gpu_backward(to_z(d["y"], z), d["angles"], SY, SX, z), z)
    assert_close(xg, d["ATy"], "AT")


def test_weight_sums_partition_of_unity():
    # examples/WoxelWeight.ipynb: per-voxel weight sums for a 9x10 grid,
    # 111 angles; the reference reports 1e-7..1e-6 errors.
    d = np.load(os.path.join(GOLDEN, "weight_sums_9x10_a111.npz"))
    SY, SX, H, A, PX = (int(v) for v in d["geom"])
    x = np.zeros((SY * SX, SY, SX, 1), np.float32)
    for i in range(SY * SX):
        x[i, i // SX, i % SX, 0] = 1
    # stack the impulses along H: one forward call covers all 90 voxels
    xs = np.ascontiguousarray(x[..., 0].transpose(1, 2, 0))
    y = gpu_forward(xs, d["angles"], PX, 0)           # (A, PX, 90)
    sums = y.astype(np.float64).sum(axis=1).T.reshape(SY, SX, A)
    err = np.abs(sums - d["sums"]).max()
    print(f"weight-sum max error {err:.2e}", flush=True)
    assert err < 1e-6


# ---------------------------------------------------------------- oracle at sizes
CASES = [
    # SY, SX, H, A, PX, angle kind
    (128, 128, 128, 192, 128, "linspace"),  # plain operator, small config
    (33, 20, 70, 45, 40, "linspace"),       # ragged H, SX != SY, odd
    (48, 48, 8, 600, 48, "linspace_open"),  # > 512 angles
    (40, 40, 64, 64, 25, "random"),         # truncated detector, unsorted angles
    (31, 37, 5, 50, 80, "wide"),            # detector wider than the diagonal, angles outside [0, pi)
    (1, 1, 1, 7, 3, "linspace"),            # single voxel
    (16, 16, 3, 1, 1, "linspace"),          # single pixel, single angle
    (40, 36, 131, 30, 44, "random"),        # 4 slices per lane, H % 4 != 0 (element DMA), ragged last block
    (24, 30, 200, 25, 30, "linspace"),      # 4 slices per lane, 16-byte DMA, idle lanes past H
    (1, 1, 257, 7, 3, "random"),            # 4 slices per lane: single voxel, two slice blocks
    (1, 9, 300, 5, 11, "linspace"),         # 4 slices per lane: one row (z_order 2 AT writes single floats)
    (13, 1, 260, 9, 1, "random"),           # 4 slices per lane: one column, one-pixel detector
    (3, 64, 256, 16, 70, "wide"),           # 4 slices per lane: z_order 2 AT on whole 16-byte tile rows
]


def make_angles(A, kind, seed=0):
    if kind == "linspace":
        return np.linspace(0, np.pi, A).astype(np.float32)
    if kind == "linspace_open":
        return np.linspace(0, np.pi, A, endpoint=False).astype(np.float32)
    rng = np.random.default_rng(seed)
    if kind == "random":
        return rng.uniform(0, np.pi, A).astype(np.float32)
    return rng.uniform(-2 * np.pi, 3 * np.pi, A).astype(np.float32)


@pytest.mark.parametrize("z", [0, 2])
@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}x{c[1]}x{c[2]}_a{c[3]}_px{c[4]}" for c in CASES])
def test_against_oracle(case, z):
    SY, SX, H, A, PX, kind = case
    print(f"oracle case {case} z{z}", flush=True)
    ang = make_angles(A, kind)
    x = seeded((SY, SX, H), 1)
    y = seeded((A, PX, H), 2)
    yo = vpo.forward_f64(x, ang, PX, 0)
    xo = vpo.backward_f64(y, ang, SY, SX, 0)
    yg = from_z(gpu_forward(to_z(x, z), ang, PX, z), z)
    xg = from_z(gpu_backward(to_z(y, z), ang, SY, SX, z), z)
    assert_close(yg, yo, "A")
    assert_close(xg, xo, "AT")
    gap = abs(np.vdot(yg.astype(np.float64), y) - np.vdot(x.astype(np.float64), xg))
    gap /= np.linalg.norm(yg) * np.linalg.norm(y)
    print(f"adjoint gap {gap:.2e}", flush=True)
    assert gap <= ADJ


# ---------------------------------------------------------------- headline sizes
@pytest.mark.parametrize("z", [0, 2])
def test_headline_512(z):
    """512^3 x 192 (PX = 512): oracle on a subset of slices + full-size properties."""
    N, A = 512, 192
    print(f"headline 512^3 z{z}", flush=True)
    ang_np = np.linspace(0, np.pi, A).astype(np.float32)
    ang = torch.tensor(ang_np, device=DEV)
    g = torch.Generator(device=DEV).manual_seed(7)
    if z == 0:
        x = torch.randn((N, N, N), device=DEV, generator=g)
        yb = torch.randn((A, N, N), device=DEV, generator=g)
        y = torch.zeros((A, N, N), device=DEV)
        xb = torch.zeros((N, N, N), device=DEV)
        _CU.forward_z0(x, y, ang)
        _CU.backward_z0(yb, xb, ang)
    else:
        x = torch.randn((N, N, N), device=DEV, generator=g)
        yb = torch.randn((N, A, N), device=DEV, generator=g)
        y = torch.zeros((N, A, N), device=DEV)
        xb = torch.zeros((N, N, N), device=DEV)
        _CU.forward_z2(x, y, ang)
        _CU.backward_z2(yb, xb, ang)
    torch.cuda.synchronize()
    # adjoint identity over the whole problem (FP64 dot products on device)
    lhs = torch.dot(y.double().flatten(), yb.double().flatten()).item()
    rhs = torch.dot(x.double().flatten(), xb.double().flatten()).item()
    gap = abs(lhs - rhs) / (y.double().norm().item() * yb.double().norm().item())
    print(f"full-size adjoint gap {gap:.2e}", flush=True)
    assert gap <= ADJ
    # oracle on slices spread over H
    hs = [0, 131, 257, 511]
    for h in hs:
        if z == 0:
            xs = x[:, :, h:h + 1].cpu().numpy()
            ys = yb[:, :, h:h + 1].cpu().numpy()
            yg = y[:, :, h:h + 1].cpu().numpy()
            xg = xb[:, :, h:h + 1].cpu().numpy()
        else:
            xs = x[h].cpu().numpy()[..., None]
            ys = yb[h].cpu().numpy()[..., None]
            yg = y[h].cpu().numpy()[..., None]
            xg = xb[h].cpu().numpy()[..., None]
        assert_close(yg, vpo.forward_f64(xs, ang_np, N, 0), f"A slice {h}")
        assert_close(xg, vpo.backward_f64(ys, ang_np, N, N, 0), f"AT slice {h}")


def test_z_orders_agree_at_512():
    N, A = 512, 192
    ang = torch.tensor(np.linspace(0, np.pi, A).astype(np.float32), device=DEV)
    g = torch.Generator(device=DEV).manual_seed(9)
    x0 = torch.randn((N, N, N), device=DEV, generator=g)
    y0 = torch.zeros((A, N, N), device=DEV)
    _CU.forward_z0(x0, y0, ang)
    x2 = x0.permute(2, 0, 1).contiguous()
    y2 = torch.zeros((N, A, N), device=DEV)
    _CU.forward_z2(x2, y2, ang)
    d = (y2 - y0.permute(2, 0, 1)).abs().max().item() / y0.abs().max().item()
    print(f"z0/z2 A max-rel difference {d:.2e}", flush=True)
    assert d <= MAX_REL
    xb0 = torch.zeros((N, N, N), device=DEV)
    _CU.backward_z0(y0, xb0, ang)
    xb2 = torch.zeros((N, N, N), device=DEV)
    _CU.backward_z2(y2, xb2, ang)
    d = (xb2 - xb0.permute(2, 0, 1)).abs().max().item() / xb0.abs().max().item()
    print(f"z0/z2 AT max-rel difference {d:.2e}", flush=True)
    assert d <= MAX_REL


@pytest.mark.parametrize("H", [5, 8])
def test_z2_many_transpose_tiles(H):
    """z_order 2 goes through layout transposes; with SY * SX = 2048^2 their
    tile count along one extent is 65536, beyond a 2-D grid's y limit.  H = 8
    takes the 16-byte path, H = 5 the element path; both equal z_order 0 bit
    for bit (the transposes are exact)."""
    SY = SX = 2048
    A, PX = 3, 24
    ang = torch.tensor(np.array([0.1, 1.2, 2.9], np.float32), device=DEV)
    g = torch.Generator(device=DEV).manual_seed(10)
    x0 = torch.randn((SY, SX, H), device=DEV, generator=g)
    y0 = torch.randn((A, PX, H), device=DEV, generator=g)
    ya = torch.zeros((A, PX, H), device=DEV)
    _CU.forward_z0(x0, ya, ang)
    y2 = torch.zeros((H, A, PX), device=DEV)
    _CU.forward_z2(x0.permute(2, 0, 1).contiguous(), y2, ang)
    assert torch.equal(y2, ya.permute(2, 0, 1))
    xa = torch.zeros((SY, SX, H), device=DEV)
    _CU.backward_z0(y0, xa, ang)
    x2 = torch.ones((H, SY, SX), device=DEV)
    _CU.backward_z2(y0.permute(2, 0, 1).contiguous(), x2, ang)
    assert torch.equal(x2, xa.permute(2, 0, 1) + 1.0)


def test_1800_angles_1024():
    """1024 x 1024 x 512 volume, 1800 angles in [0, pi): angle chunking past 512."""
    SY = SX = 1024
    H, A = 512, 1800
    print("1024^2 x 512 x 1800", flush=True)
    ang_np = np.linspace(0, np.pi, A, endpoint=False).astype(np.float32)
    ang = torch.tensor(ang_np, device=DEV)
    g = torch.Generator(device=DEV).manual_seed(11)
    x = torch.randn((SY, SX, H), device=DEV, generator=g)
    y = torch.zeros((A, SX, H), device=DEV)
    _CU.forward_z0(x, y, ang)
    xb = torch.zeros((SY, SX, H), device=DEV)
    _CU.backward_z0(y, xb, ang)
    torch.cuda.synchronize()
    for h in (3, 400):
        xs = x[:, :, h:h + 1].cpu().numpy()
        yo = vpo.forward_f64(xs, ang_np, SX, 0)
        assert_close(y[:, :, h:h + 1].cpu().numpy(), yo, f"A slice {h}")
        xo = vpo.backward_f64(y[:, :, h:h + 1].cpu().numpy(), ang_np, SY, SX, 0)
        assert_close(xb[:, :, h:h + 1].cpu().numpy(), xo, f"AT slice {h}")


# ---------------------------------------------------------------- semantics
@pytest.mark.parametrize("z", [0, 2])
def test_accumulates_and_is_deterministic(z):
    SY, SX, H, A, PX = 40, 36, 70, 50, 52
    ang = torch.tensor(make_angles(A, "linspace"), device=DEV)
    g = torch.Generator(device=DEV).manual_seed(3)
    xs = (SY, SX, H) if z == 0 else (H, SY, SX)
    ys = (A, PX, H) if z == 0 else (H, A, PX)
    fwd = _CU.forward_z0 if z == 0 else _CU.forward_z2
    bwd = _CU.backward_z0 if z == 0 else _CU.backward_z2
    x = torch.randn(xs, device=DEV, generator=g)
    y = torch.randn(ys, device=DEV, generator=g)
    fresh = torch.zeros(ys, device=DEV)
    fwd(x, fresh, ang)
    again = torch.zeros(ys, device=DEV)
    fwd(x, again, ang)
    assert torch.equal(fresh, again)            # bitwise deterministic
    acc = y.clone()
    fwd(x, acc, ang)
    assert torch.equal(acc, y + fresh)          # out += A x
    fb = torch.zeros(xs, device=DEV)
    bwd(y, fb, ang)
    fb2 = torch.zeros(xs, device=DEV)
    bwd(y, fb2, ang)
    assert torch.equal(fb, fb2)
    accb = x.clone()
    bwd(y, accb, ang)
    assert torch.equal(accb, x + fb)            # out += AT y


def test_cu_validation_on_device():
    ang = torch.zeros(7, device=DEV)
    x = torch.zeros((6, 5, 3), device=DEV)
    y = torch.zeros((7, 11, 3), device=DEV)
    with pytest.raises(RuntimeError, match="contiguous"):
        _CU.forward_z0(x.transpose(0, 1), y, ang)
    with pytest.raises(RuntimeError, match="float32"):
        _CU.forward_z0(x.double(), y, ang)
    with pytest.raises(RuntimeError, match="3D"):
        _CU.forward_z0(x[0], y, ang)
    with pytest.raises(RuntimeError, match="angle"):
        _CU.forward_z0(x, y, torch.zeros(6, device=DEV))
    with pytest.raises(RuntimeError, match="H mismatch"):
        _CU.backward_z0(y, torch.zeros((6, 5, 4), device=DEV), ang)
