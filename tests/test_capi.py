"""The C-ABI library loads and exports every symbol declared in include/*.h.

Only argument validation is exercised here (it runs on the host before any
HIP call), so these tests need no GPU.
"""
import ctypes
import glob
import os
import re

import pytest

from tests.conftest import ROOT

import voxelproj_amd
from voxelproj_amd import _lib


def _declared_symbols():
    names = set()
    for hdr in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(hdr).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"VP_API\s+[\w\s\*]+?\b(vp_\w+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_header_declares_entry_points():
    names = _declared_symbols()
    assert {"vp_forward", "vp_backward", "vp_last_error", "vp_abi_version"} <= names


def test_library_exports_declared_symbols():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in _declared_symbols():
        assert hasattr(lib, name), name
    assert set(_lib.EXPORTS) <= _declared_symbols()


def test_abi_version():
    assert voxelproj_amd.load().vp_abi_version() == _lib.ABI_VERSION


def _call(fn, z=0, ptr=1, A=3, SY=4, SX=4, H=2, PX=5):
    lib = voxelproj_amd.load()
    p = ctypes.c_void_p(ptr) if ptr else None
    return getattr(lib, fn)(z, p, p, p, A, SY, SX, H, PX, None)


@pytest.mark.parametrize("fn", ["vp_forward", "vp_backward"])
def test_bad_z_order_rejected(fn):
    assert _call(fn, z=1) == -1
    assert b"z_order" in voxelproj_amd.load().vp_last_error()


@pytest.mark.parametrize("fn", ["vp_forward", "vp_backward"])
def test_null_pointer_rejected(fn):
    assert _call(fn, ptr=0) == -1
    assert b"null" in voxelproj_amd.load().vp_last_error()


@pytest.mark.parametrize("fn", ["vp_forward", "vp_backward"])
def test_negative_and_oversized_extents_rejected(fn):
    assert _call(fn, SX=-1) == -1
    assert _call(fn, PX=40000) == -2
    assert _call(fn, A=70000) == -2


@pytest.mark.parametrize("fn", ["vp_forward", "vp_backward"])
def test_empty_problem_is_a_no_op(fn):
    # zero extents return success without touching the (null) pointers
    assert _call(fn, ptr=0, A=0) == 0
    assert _call(fn, ptr=0, H=0) == 0
    assert voxelproj_amd.load().vp_last_error() == b""


def _call_geom(fn, z=0, ptr=1, A=3, SY=4, SX=4, H=2, PX=5, du=None, dv=None, du_max=0.0):
    lib = voxelproj_amd.load()
    p = ctypes.c_void_p(ptr) if ptr else None
    return getattr(lib, fn)(z, p, p, p, A, SY, SX, H, PX, du, dv, du_max, None)


@pytest.mark.parametrize("fn", ["vp_forward_geom", "vp_backward_geom", "vp_forward_geom_f64",
                                "vp_backward_geom_f64"])
def test_geometry_entry_points_validate(fn):
    lib = voxelproj_amd.load()
    assert _call_geom(fn, z=1) == -1 and b"z_order" in lib.vp_last_error()
    assert _call_geom(fn, ptr=0) == -1 and b"null" in lib.vp_last_error()
    assert _call_geom(fn, PX=40000) == -2
    # du_max must be a finite, non-negative bound
    assert _call_geom(fn, du=ctypes.c_void_p(1), du_max=-1.0) == -1
    assert _call_geom(fn, du=ctypes.c_void_p(1), du_max=float("nan")) == -1
    assert b"du_max" in lib.vp_last_error()
    assert _call_geom(fn, ptr=0, A=0) == 0


def test_vshift_validates():
    lib = voxelproj_amd.load()
    p = ctypes.c_void_p(1)
    assert lib.vp_vshift(1, 0, p, p, 3, 5, 4, 0, 4, 0, p, 0, None) == -1
    assert lib.vp_vshift(0, 2, p, p, 3, 5, 4, 0, 4, 0, p, 0, None) == -1
    assert lib.vp_vshift(0, 0, p, p, 3, 5, 4, 0, 4, 0, p, 2, None) == -1
    assert lib.vp_vshift(0, 0, p, None, 3, 5, 4, 0, 4, 0, p, 0, None) == -1
    assert lib.vp_vshift(0, 0, p, p, 3, -5, 4, 0, 4, 0, p, 0, None) == -1
    assert lib.vp_vshift(0, 0, None, None, 3, 5, 4, 0, 0, 0, None, 0, None) == 0  # empty output


class _BatchItem(ctypes.Structure):
    _fields_ = [("inp", ctypes.c_void_p), ("out", ctypes.c_void_p), ("angles", ctypes.c_void_p),
                ("n_angles", ctypes.c_int32), ("SY", ctypes.c_int32), ("SX", ctypes.c_int32),
                ("H", ctypes.c_int32), ("PX", ctypes.c_int32), ("accumulate", ctypes.c_int32)]


def test_batched_forward_offset_span_limit():
    """Forward variants 1 and 2 stage several voxels per DMA instruction through
    one 32-bit buffer offset; items whose SX * H exceeds that are rejected
    (variant 4, one voxel per instruction, takes them)."""
    lib = voxelproj_amd.load()
    big = (_BatchItem * 1)(_BatchItem(1, 1, 1, 1, 1, 32768, 16384, 1, 0))
    for v in (1, 2):
        assert lib.vp_forward_batched(v, big, big, 1, 0, None) == -2
        assert b"too large for variant" in lib.vp_last_error()
    bad = (_BatchItem * 1)(_BatchItem(1, 1, 1, 1, 1, 4, 2, 1, 5))
    assert lib.vp_forward_batched(1, bad, bad, 1, 0, None) == -1  # accumulate not 0 / 1


def test_workspace_sizes():
    lib = voxelproj_amd.load()
    A, SY, SX, H, PX = 7, 6, 5, 3, 11
    nx, ny = SY * SX * H, A * PX * H
    assert lib.vp_workspace_elems(0, 0, A, SY, SX, H, PX, 0) == 0
    assert lib.vp_workspace_elems(0, 2, A, SY, SX, H, PX, 0) == nx + ny
    assert lib.vp_workspace_elems(0, 0, A, SY, SX, H, PX, 1) == ny
    assert lib.vp_workspace_elems(0, 2, A, SY, SX, H, PX, 1) == nx + 2 * ny
    assert lib.vp_workspace_elems(1, 2, A, SY, SX, H, PX, 0) == 0
    assert lib.vp_workspace_elems(1, 0, A, SY, SX, H, PX, 1) == ny
    assert lib.vp_workspace_elems(0, 1, A, SY, SX, H, PX, 0) < 0


def test_caller_workspace_is_validated_before_device_work():
    lib = voxelproj_amd.load()
    A, SY, SX, H, PX = 7, 6, 5, 3, 11
    p = ctypes.c_void_p(1)
    for fn in (lib.vp_forward_ws, lib.vp_backward_ws):
        # negative size, or a size without a buffer
        assert fn(2, p, p, p, A, SY, SX, H, PX, None, None, 0.0, p, -1, None) == -1
        assert fn(2, p, p, p, A, SY, SX, H, PX, None, None, 0.0, None, 100, None) == -1
        # z_order 2 needs at least one slice of volume and sinogram
        one = SY * SX + A * PX
        assert fn(2, p, p, p, A, SY, SX, H, PX, None, None, 0.0, p, one - 1, None) == -2
        assert b"workspace" in lib.vp_last_error()
        # with dv it needs the whole size
        need = lib.vp_workspace_elems(0, 2, A, SY, SX, H, PX, 1)
        assert fn(2, p, p, p, A, SY, SX, H, PX, None, p, 0.0, p, need - 1, None) == -2


def test_noalloc_validation_before_device_work():
    lib = voxelproj_amd.load()
    p = ctypes.c_void_p(1)
    n = lib.vp_noalloc_scratch_elems()
    assert n == 16 << 20
    # bad op, bad z_order
    assert lib.vp_project_noalloc(2, 0, p, p, p, 3, 4, 4, 2, 5, None) == -1
    assert lib.vp_project_noalloc(0, 1, p, p, p, 3, 4, 4, 2, 5, None) == -1
    # z_order 2 with one slice larger than the scratch
    side = 4097
    assert lib.vp_project_noalloc(1, 2, p, p, p, 1, side, side, 1, side, None) == -2
    assert b"scratch" in lib.vp_last_error()
    # empty problems are a no-op
    assert lib.vp_project_noalloc(0, 2, p, p, p, 0, 4, 4, 2, 5, None) == 0


class _TransposeItem(ctypes.Structure):
    _fields_ = [("inp", ctypes.c_void_p), ("out", ctypes.c_void_p), ("rows", ctypes.c_int64),
                ("cols", ctypes.c_int64), ("accumulate", ctypes.c_int32), ("reserved", ctypes.c_int32)]


def test_transpose_batched_validates_before_device_work():
    lib = voxelproj_amd.load()
    p = ctypes.c_void_p(1)
    ok = (_TransposeItem * 1)(_TransposeItem(1, 1, 3, 4, 0, 0))
    assert lib.vp_transpose_batched(ok, ok, -1, 0, None) == -1       # negative count
    assert lib.vp_transpose_batched(None, ok, 1, 0, None) == -1      # no host descriptors
    assert lib.vp_transpose_batched(ok, None, 1, 0, None) == -1      # no device descriptors
    assert lib.vp_transpose_batched(None, None, 0, 0, None) == 0     # nothing to do
    for bad in (_TransposeItem(1, 1, -3, 4, 0, 0), _TransposeItem(None, 1, 3, 4, 0, 0),
                _TransposeItem(1, 1, 3, 4, 2, 0), _TransposeItem(1, 1, 3, 4, 0, 7)):
        arr = (_TransposeItem * 2)(_TransposeItem(1, 1, 2, 2, 1, 0), bad)
        assert lib.vp_transpose_batched(arr, p, 2, 0, None) == -1
        assert b"transpose item 1" in lib.vp_last_error()
    huge = (_TransposeItem * 1)(_TransposeItem(1, 1, 1 << 30, 1 << 20, 0, 0))
    assert lib.vp_transpose_batched(huge, p, 1, 0, None) == -2


def test_launch_variant_rule_without_a_device():
    """vp_launch_variant: the variant a single launch uses (256 compute units
    assumed without a device).  A above 128 slices takes the strip width whose
    grid fills the device best (rounds of workgroups x (P + 4)); below, the
    H-based default, 2 slices per lane dropping to 1 on small grids; AT always
    the H-based default."""
    lib = voxelproj_amd.load()
    lv = lib.vp_launch_variant
    # (n_angles, SY, SX, H, PX) -> variant, worked out by hand for 256 CUs:
    # 256^3: 12 x 16 x 1 = 192 blocks of 22 pixels (26) vs 16 x 16 = 256 of 16 (20)
    assert lv(0, 192, 256, 256, 256, 256) == 416
    # 288^3: 2 slice blocks; 16 strips of 18 pixels x 16 x 2 = 512 = two rounds (2 x 22)
    assert lv(0, 192, 288, 288, 288, 288) == 418
    # 320^3: 16 strips of 20 pixels x 32 = 512 (2 x 24) against 15 of 22 x 32 = 480 (2 x 26)
    assert lv(0, 192, 320, 320, 320, 320) == 420
    # 384^3: 24 strips of 16 x 32 = 768, three rounds (3 x 20) against 18 of 22 (3 x 26)
    assert lv(0, 192, 384, 384, 384, 384) == 416
    # 512^3: 24 strips of 22 x 32 = 768 (3 x 26) against 32 of 16 x 32 = 1024 (4 x 20)
    assert lv(0, 192, 512, 512, 512, 512) == 4
    assert lv(0, 1800, 1024, 1024, 512, 1024) == 4
    # at most 128 slices: the H-based default, 2 -> 1 on grids smaller than the device
    assert lv(0, 192, 128, 128, 128, 128) == 1
    assert lv(0, 1800, 512, 512, 128, 512) == 2
    assert lv(0, 360, 256, 256, 32, 256) == 1
    for H, v in ((32, 32), (64, 64), (128, 128), (129, 4), (512, 4)):
        assert lv(1, 192, 256, 256, H, 256) == v == lib.vp_default_variant(1, H)
    assert lv(2, 192, 64, 64, 64, 64) == -1
    assert lv(0, 0, 64, 64, 64, 64) == -1
    assert lv(0, 10, 64, 64, -1, 64) == -1
