"""Pins of the translated-geometry oracle (oracle.forward_geom_f64 /
backward_geom_f64), CPU only.

The translations are an extension (the reference lists them as future work,
README.md:65-68), so no reference output exists; the oracle is pinned by:
* du = dv = 0 reproduces the untranslated oracle exactly;
* integer du is an exact pixel shift of the untranslated sinogram (taken on a
  wider detector so that nothing is cut off), integer dv an exact H shift;
* fractional dv is the linear interpolation of two integer shifts;
* fractional du against the polygon-clipping brute force (tests/_brute.py,
  areas of the unit voxel inside the shifted detector strips);
* an integer in-plane translation of the volume equals du = -(dx cos + dy sin);
* the adjoint identity with per-angle du and dv.
"""
import numpy as np
import pytest

import oracle as vpo
from tests import _brute

rng = np.random.default_rng(7)


def _geom(SY=9, SX=11, H=6, A=13, PX=14):
    ang = np.linspace(0, np.pi, A, endpoint=False).astype(np.float32)
    return SY, SX, H, A, PX, ang


@pytest.mark.parametrize("z_order", [0, 2])
def test_zero_translation_is_plain(z_order):
    SY, SX, H, A, PX, ang = _geom()
    x = rng.standard_normal(vpo.x_shape(z_order, SY, SX, H))
    y = rng.standard_normal(vpo.y_shape(z_order, A, PX, H))
    z = np.zeros(A)
    assert np.array_equal(vpo.forward_geom_f64(x, ang, PX, z_order, z, z), vpo.forward_f64(x, ang, PX, z_order))
    assert np.array_equal(vpo.backward_geom_f64(y, ang, SY, SX, z_order, z, z),
                          vpo.backward_f64(y, ang, SY, SX, z_order))


@pytest.mark.parametrize("k", [1, 3, -2])
def test_integer_du_is_pixel_shift(k):
    SY, SX, H, A, PX, ang = _geom()
    x = rng.standard_normal((SY, SX, H))
    got = vpo.forward_geom_f64(x, ang, PX, 0, du=k)
    m = abs(k)
    wide = vpo.forward_f64(x, ang, PX + 2 * m, 0)  # pc + m: pixel p of got is pixel p + m + k of wide
    assert np.allclose(got, wide[:, m + k:m + k + PX, :], rtol=0, atol=1e-13)


@pytest.mark.parametrize("k", [1, -2, 5])
def test_integer_dv_is_slice_shift(k):
    SY, SX, H, A, PX, ang = _geom(H=7)
    x = rng.standard_normal((SY, SX, H))
    got = vpo.forward_geom_f64(x, ang, PX, 0, dv=k)
    a0 = vpo.forward_f64(x, ang, PX, 0)
    ref = np.zeros_like(a0)
    for h in range(H):
        if 0 <= h + k < H:
            ref[:, :, h] = a0[:, :, h + k]
    assert np.array_equal(got, ref)


def test_fractional_dv_is_interpolation():
    SY, SX, H, A, PX, ang = _geom(H=8)
    x = rng.standard_normal((SY, SX, H))
    dv = np.float32(-1.3)
    k, f = int(np.floor(dv)), float(dv) - np.floor(float(dv))
    got = vpo.forward_geom_f64(x, ang, PX, 0, dv=dv)
    ref = (1 - f) * vpo.forward_geom_f64(x, ang, PX, 0, dv=k) + f * vpo.forward_geom_f64(x, ang, PX, 0, dv=k + 1)
    assert np.allclose(got, ref, rtol=0, atol=1e-12)


@pytest.mark.parametrize("phi", [0.0, 0.3, np.pi / 4, 1.2, 2.5])
@pytest.mark.parametrize("du", [0.37, -1.61, 2.5])
def test_fractional_du_matches_polygon_brute_force(phi, du):
    SY, SX, PX = 5, 6, 9
    ang = np.array([phi], np.float32)
    W = _brute.weights(phi, SY, SX, PX, du=du)  # (PX, SY, SX)
    eye = np.eye(SY * SX).reshape(SY * SX, SY, SX)
    # column v of A_du is the sinogram of a unit impulse at voxel v
    x = eye.transpose(1, 2, 0)  # (SY, SX, H = SY*SX): slice v holds impulse v
    y = vpo.forward_geom_f64(x, ang, PX, 0, du=du)  # (1, PX, SY*SX)
    assert np.allclose(y[0], W.reshape(PX, SY * SX), rtol=0, atol=1e-12)


def test_volume_translation_equals_du():
    SY, SX, H, A, PX, ang = _geom(SY=12, SX=12, H=3, A=11, PX=20)
    x = np.zeros((SY, SX, H))
    x[3:8, 2:9] = rng.standard_normal((5, 7, H))  # zero border, room to move
    dx, dy = 2, -1
    xs = np.zeros_like(x)
    xs[3 + dy:8 + dy, 2 + dx:9 + dx] = x[3:8, 2:9]  # content moved by (+dx, +dy)
    du = -(dx * np.cos(ang.astype(np.float64)) + dy * np.sin(ang.astype(np.float64)))
    got = vpo.forward_geom_f64(x, ang, PX, 0, du=du)
    ref = vpo.forward_f64(xs, ang, PX, 0)
    # du is rounded to float32 on the way in: positions move by <= 1e-7 px
    assert np.abs(got - ref).max() <= 1e-6 * np.abs(ref).max()


@pytest.mark.parametrize("z_order", [0, 2])
def test_translated_pair_is_adjoint(z_order):
    SY, SX, H, A, PX, ang = _geom(H=9, A=17)
    du = rng.uniform(-4, 4, A)
    dv = rng.uniform(-3, 3, A)
    x = rng.standard_normal(vpo.x_shape(z_order, SY, SX, H))
    y = rng.standard_normal(vpo.y_shape(z_order, A, PX, H))
    ax = vpo.forward_geom_f64(x, ang, PX, z_order, du, dv)
    aty = vpo.backward_geom_f64(y, ang, SY, SX, z_order, du, dv)
    gap = abs(np.vdot(ax, y) - np.vdot(x, aty)) / (np.linalg.norm(ax) * np.linalg.norm(y))
    assert gap < 1e-14
