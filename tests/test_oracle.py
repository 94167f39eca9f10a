"""Pins the CPU oracle (test infrastructure) before it is trusted as a checker.

Each property below is chosen so that a plausible mistake in the oracle -- a
wrong footprint half-width, a swapped sin/cos, an off-by-one centre, a
transposed layout, a dropped tap -- fails at least one of them:

* brute force: polygon-clipping areas (tests/_brute.py) vs the closed-form
  footprint, all (voxel, pixel) pairs, several angles incl. 0, pi/4, pi/2;
* closed forms: phi = 0 gives column sums; pi/4 gives the triangle weights
  sqrt2 - 1/2 and (3/2 - sqrt2)/2;
* invariants: partition of unity (README.md:13-15, examples/WoxelWeight.ipynb),
  AT(1) = n_angles, phi -> phi + pi mirrors the detector, linearity,
  z_order 0/2 equivalence, H-slice independence;
* independent code paths: the ray-driven FP32 transcription of the reference
  forward (sp_forward_z0.cuh) against the voxel-driven FP64 oracle, and the
  adjoint identity between them;
* the reference's own tests, restated (test/test_adjoitness.py:20-51).
"""
import numpy as np
import pytest

import oracle as vpo
from tests import _brute


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b))


@pytest.mark.parametrize("phi", [0.0, np.pi / 4, np.pi / 2, 0.3, 1.1, 2.0, 2.9, np.pi])
@pytest.mark.parametrize("geom", [(6, 5, 11), (4, 7, 4), (5, 5, 9)])
def test_weights_match_polygon_brute_force(phi, geom):
    SY, SX, PX = geom
    W = _brute.weights(phi, SY, SX, PX)
    rng = np.random.default_rng(3)
    x = rng.standard_normal((SY, SX, 2)).astype(np.float32)
    yb = np.einsum("pyx,yxh->ph", W, x.astype(np.float64))
    yo = vpo.forward_f64(x, [phi], PX, 0)[0]
    np.testing.assert_allclose(yo, yb, rtol=0, atol=1e-12)


def test_brute_force_single_weights():
    # A unit impulse projected through the oracle equals the polygon areas.
    SY, SX, PX = 3, 4, 7
    for phi in (0.2, 0.7, 1.3):
        W = _brute.weights(phi, SY, SX, PX)
        for iy in range(SY):
            for ix in range(SX):
                x = np.zeros((SY, SX, 1), np.float32)
                x[iy, ix, 0] = 1
                y = vpo.forward_f64(x, [phi], PX, 0)[0, :, 0]
                np.testing.assert_allclose(y, W[:, iy, ix], atol=1e-13)


def test_phi_zero_is_column_sum():
    rng = np.random.default_rng(0)
    SY, SX, H = 7, 9, 3
    x = rng.standard_normal((SY, SX, H)).astype(np.float32)
    y = vpo.forward_f64(x, [0.0], SX, 0)[0]
    np.testing.assert_allclose(y, x.astype(np.float64).sum(0), atol=1e-12)


def test_phi_quarter_pi_triangle():
    # centred voxel, phi = pi/4: the footprint is a triangle of half-base
    # sqrt2/2 and height sqrt2; the central bin holds sqrt2 - 1/2.
    x = np.zeros((1, 1, 1), np.float32)
    x[0, 0, 0] = 1
    y = vpo.forward_f64(x, [np.pi / 4], 3, 0)[0, :, 0]
    side = (1.5 - np.sqrt(2)) / 2
    np.testing.assert_allclose(y, [side, np.sqrt(2) - 0.5, side], atol=2e-7)


@pytest.mark.parametrize("z_order", [0, 2])
def test_partition_of_unity(z_order):
    rng = np.random.default_rng(1)
    SY, SX, H, A = 9, 10, 2, 111
    PX = SX + SY
    ang = np.linspace(0, np.pi, A).astype(np.float32)
    x = rng.standard_normal(vpo.x_shape(z_order, SY, SX, H)).astype(np.float32)
    y = vpo.forward_f64(x, ang, PX, z_order)
    if z_order == 0:
        tot = y.sum(axis=1)            # (A, H)
        ref = x.astype(np.float64).sum(axis=(0, 1))[None, :]
    else:
        tot = y.sum(axis=2).T          # (A, H)
        ref = x.astype(np.float64).sum(axis=(1, 2))[None, :]
    np.testing.assert_allclose(tot, np.broadcast_to(ref, tot.shape), atol=1e-11)


def test_backprojection_of_ones_counts_angles():
    SY, SX, H, A = 8, 6, 2, 37
    PX = SX + SY
    ang = np.linspace(0, np.pi, A, endpoint=False).astype(np.float32)
    y = np.ones((A, PX, H), np.float32)
    x = vpo.backward_f64(y, ang, SY, SX, 0)
    np.testing.assert_allclose(x, A, atol=1e-11)


def test_mirror_by_pi():
    rng = np.random.default_rng(2)
    SY, SX, H, PX = 7, 8, 2, 13
    x = rng.standard_normal((SY, SX, H)).astype(np.float32)
    for phi in (0.1, 0.8, 1.9):
        y1 = vpo.forward_f64(x, [phi], PX, 0)[0]
        y2 = vpo.forward_f64(x, [np.float32(phi) + np.float32(np.pi)], PX, 0)[0]
        np.testing.assert_allclose(y2, y1[::-1], atol=5e-6)


def test_linearity_and_slice_independence():
    rng = np.random.default_rng(4)
    SY, SX, H, A, PX = 6, 7, 4, 9, 10
    ang = rng.uniform(0, np.pi, A).astype(np.float32)
    x1 = rng.standard_normal((SY, SX, H)).astype(np.float64)
    x2 = rng.standard_normal((SY, SX, H)).astype(np.float64)
    y = vpo.forward_f64(2.0 * x1 - 3.0 * x2, ang, PX, 0)
    y1 = vpo.forward_f64(x1, ang, PX, 0)
    y2 = vpo.forward_f64(x2, ang, PX, 0)
    np.testing.assert_allclose(y, 2 * y1 - 3 * y2, atol=1e-12)
    part = vpo.forward_f64(x1, ang, PX, 0, h0=1, h1=3)
    np.testing.assert_allclose(part[..., 1:3], y1[..., 1:3], atol=0)
    assert np.all(part[..., 0] == 0) and np.all(part[..., 3] == 0)


@pytest.mark.parametrize("op", ["forward", "backward"])
def test_z_orders_agree(op):
    rng = np.random.default_rng(5)
    SY, SX, H, A, PX = 6, 9, 3, 11, 12
    ang = np.linspace(0, np.pi, A).astype(np.float32)
    if op == "forward":
        x0 = rng.standard_normal((SY, SX, H)).astype(np.float32)
        y0 = vpo.forward_f64(x0, ang, PX, 0)
        y2 = vpo.forward_f64(np.ascontiguousarray(x0.transpose(2, 0, 1)), ang, PX, 2)
        np.testing.assert_array_equal(y2, y0.transpose(2, 0, 1))
    else:
        y0 = rng.standard_normal((A, PX, H)).astype(np.float32)
        x0 = vpo.backward_f64(y0, ang, SY, SX, 0)
        x2 = vpo.backward_f64(np.ascontiguousarray(y0.transpose(2, 0, 1)), ang, SY, SX, 2)
        np.testing.assert_array_equal(x2, x0.transpose(2, 0, 1))


@pytest.mark.parametrize("PX", [11, 4, 30])
def test_ray_driven_transcription_matches_oracle(PX):
    # Independent algorithm (ray traversal with box clipping, as in the
    # reference forward kernel) agrees with the voxel-driven FP64 oracle.
    rng = np.random.default_rng(6)
    SY, SX, H, A = 6, 5, 3, 31
    ang = np.linspace(0, np.pi, A).astype(np.float32)
    x = rng.standard_normal((SY, SX, H)).astype(np.float32)
    y64 = vpo.forward_f64(x, ang, PX, 0)
    y32 = vpo.ref_forward_f32(x, ang, PX, 0)
    assert rel(y32, y64) < 1e-6
    yv = rng.standard_normal((A, PX, H)).astype(np.float32)
    x64 = vpo.backward_f64(yv, ang, SY, SX, 0)
    x32 = vpo.ref_backward_f32(yv, ang, SY, SX, 0)
    assert rel(x32, x64) < 1e-6
    # adjoint identity across the two independent code paths
    lhs = np.vdot(y32.astype(np.float64), yv)
    rhs = np.vdot(x.astype(np.float64), x64)
    assert abs(lhs - rhs) / (np.linalg.norm(y32) * np.linalg.norm(yv)) < 1e-6


@pytest.mark.parametrize("z_order", [0, 2])
@pytest.mark.parametrize("detector", ["full", "interior"])
def test_reference_adjointness_restated(z_order, detector):
    # test/test_adjoitness.py:12-51, seeded (the reference's SEED is unused there).
    SX, SY, H, n_angles = 5, 6, 3, 31
    PX = SX + SY if detector == "full" else min(SY, SX) - 1
    angles = np.linspace(0, np.pi, n_angles).astype(np.float32)
    rng = np.random.default_rng(128437)
    x = rng.random(vpo.x_shape(z_order, SY, SX, H)).astype(np.float32)
    y = rng.random(vpo.y_shape(z_order, n_angles, PX, H)).astype(np.float32)
    for fwd, bwd in ((vpo.forward_f64, vpo.backward_f64), (vpo.ref_forward_f32, vpo.ref_backward_f32)):
        Ax = fwd(x, angles, PX, z_order)
        ATy = bwd(y, angles, SY, SX, z_order)
        lhs = np.vdot(Ax.astype(np.float64), y)
        rhs = np.vdot(x.astype(np.float64), ATy.astype(np.float64))
        np.testing.assert_allclose(lhs, rhs, rtol=1e-6, atol=1e-6)


def test_nnz_per_voxel_angle():
    ang = np.linspace(0, np.pi, 192).astype(np.float32)
    n = vpo.nnz(ang, 128, 128, 128)
    assert abs(n / (128 * 128 * 192) - 2.1306) < 5e-4
    # exact count on a tiny case equals the dense brute-force support
    ang2 = np.array([0.3, 1.2], np.float32)
    cnt = sum(int(np.count_nonzero(_brute.weights(p, 5, 4, 8) > 1e-13)) for p in ang2)
    assert vpo.nnz(ang2, 5, 4, 8) == cnt


# ------------------------------------------------ the scalar oracle exports
def _brute_cdf(phi, t):
    """Area of the unit voxel centred at 0 whose detector coordinate
    x cos + y sin is below t (one half-plane clip of the square)."""
    c, s = np.cos(np.float64(np.float32(phi))), np.sin(np.float64(np.float32(phi)))
    sq = [(-0.5, -0.5), (0.5, -0.5), (0.5, 0.5), (-0.5, 0.5)]
    return _brute._area(_brute._clip(sq, -c, -s, -t))


@pytest.mark.parametrize("phi", [0.0, 0.2, np.pi / 4, 1.0, np.pi / 2, 2.5])
def test_cdf_matches_half_plane_area(phi):
    for t in np.linspace(-1.0, 1.0, 41):
        assert abs(vpo.cdf_f64(phi, t) - _brute_cdf(phi, t)) < 1e-13, (phi, t)


@pytest.mark.parametrize("phi", [0.0, 0.3, np.pi / 4, 1.2, 2.9])
def test_weight_matches_polygon_area(phi):
    # one voxel at the volume centre (u = pc): its weights on a 5-pixel
    # detector are the strip areas of the brute-force definition
    W = _brute.weights(phi, 1, 1, 5)[:, 0, 0]
    for p in range(5):
        assert abs(vpo.weight_f64(phi, p, 2.0) - W[p]) < 1e-13, (phi, p)
    # and W(p, u) = F(p + 1/2 - u) - F(p - 1/2 - u) for off-centre positions
    for u in (1.3, 2.45, 2.5, 3.01):
        for p in range(5):
            w = vpo.cdf_f64(phi, p + 0.5 - u) - vpo.cdf_f64(phi, p - 0.5 - u)
            assert abs(vpo.weight_f64(phi, p, u) - w) < 1e-14


def test_angle_table_describes_the_footprint():
    phis = np.array([0.0, 0.1, np.pi / 4, 1.0, np.pi / 2, 2.2, np.pi], np.float32)
    tab = vpo.angle_table_f64(phis)
    for (c, s, a, b, h), phi in zip(tab, phis):
        np.testing.assert_allclose([c, s], [np.cos(np.float64(phi)), np.sin(np.float64(phi))],
                                   rtol=0, atol=1e-15)
        # b: support half-width of F; h: its slope on the plateau [-a, a];
        # unit area: h (a + b) = 1
        assert vpo.cdf_f64(phi, b) == pytest.approx(1.0, abs=1e-14)
        assert vpo.cdf_f64(phi, -b) == pytest.approx(0.0, abs=1e-14)
        assert vpo.cdf_f64(phi, b - 1e-4) < 1.0
        eps = 1e-7
        slope = (_brute_cdf(phi, eps) - _brute_cdf(phi, -eps)) / (2 * eps)
        assert slope == pytest.approx(h, rel=1e-6)
        assert h * (a + b) == pytest.approx(1.0, abs=1e-14)
        if a > 1e-3:  # the plateau ends at +-a: F is linear up to a, not beyond
            mid = _brute_cdf(phi, a / 2)
            assert mid == pytest.approx(0.5 + h * a / 2, abs=1e-12)
