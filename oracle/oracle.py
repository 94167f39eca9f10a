"""ctypes front end of the CPU oracle library oracle/libvp_oracle.so.

TEST INFRASTRUCTURE ONLY -- see oracle/__init__.py.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libvp_oracle.so")
_lib = None

_i64 = ctypes.c_int64
_int = ctypes.c_int
_vp = ctypes.c_void_p


def build():
    """Compile oracle/libvp_oracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    lib = ctypes.CDLL(_LIB_PATH)
    sig = [_vp, _int, _vp, _vp, _int, _i64, _i64, _i64, _i64, _int, _i64, _i64, _int]
    lib.vpo_forward_f64.argtypes = sig
    lib.vpo_forward_f64.restype = _int
    lib.vpo_backward_f64.argtypes = sig
    lib.vpo_backward_f64.restype = _int
    sigg = sig[:-1] + [_vp, _vp, _int]
    lib.vpo_forward_geom_f64.argtypes = sigg
    lib.vpo_forward_geom_f64.restype = _int
    lib.vpo_backward_geom_f64.argtypes = sigg
    lib.vpo_backward_geom_f64.restype = _int
    sig32 = [_vp, _vp, _vp, _int, _i64, _i64, _i64, _i64, _int, _i64, _i64, _int]
    lib.vpo_ref_forward_f32.argtypes = sig32
    lib.vpo_ref_forward_f32.restype = _int
    lib.vpo_ref_backward_f32.argtypes = sig32
    lib.vpo_ref_backward_f32.restype = _int
    lib.vpo_nnz.argtypes = [_vp, _int, _i64, _i64, _i64, _int]
    lib.vpo_nnz.restype = _i64
    lib.vpo_weight_f64.argtypes = [ctypes.c_float, ctypes.c_double, ctypes.c_double]
    lib.vpo_weight_f64.restype = ctypes.c_double
    lib.vpo_cdf_f64.argtypes = [ctypes.c_float, ctypes.c_double]
    lib.vpo_cdf_f64.restype = ctypes.c_double
    lib.vpo_angle_table64.argtypes = [_vp, _int, _vp]
    lib.vpo_angle_table64.restype = None
    lib.vpo_max_threads.argtypes = []
    lib.vpo_max_threads.restype = _int
    _lib = lib
    return lib


def max_threads():
    return int(load().vpo_max_threads())


def x_shape(z_order, SY, SX, H):
    return {0: (SY, SX, H), 2: (H, SY, SX)}[z_order]


def y_shape(z_order, A, PX, H):
    return {0: (A, PX, H), 2: (H, A, PX)}[z_order]


def _vol_dims(x, z_order):
    if z_order == 0:
        SY, SX, H = x.shape
    elif z_order == 2:
        H, SY, SX = x.shape
    else:
        raise ValueError("z_order must be 0 or 2")
    return SY, SX, H


def _sino_dims(y, z_order):
    if z_order == 0:
        A, PX, H = y.shape
    elif z_order == 2:
        H, A, PX = y.shape
    else:
        raise ValueError("z_order must be 0 or 2")
    return A, PX, H


def _angles32(angles):
    return np.ascontiguousarray(np.asarray(angles, dtype=np.float32).reshape(-1))


def _input(a):
    a = np.asarray(a)
    if a.dtype == np.float64:
        return np.ascontiguousarray(a), 1
    return np.ascontiguousarray(a, dtype=np.float32), 0


def _hrange(H, h0, h1):
    h1 = H if h1 is None else h1
    if not (0 <= h0 <= h1 <= H):
        raise ValueError("bad H range")
    return h0, h1


def forward_f64(x, angles, PX=None, z_order=0, h0=0, h1=None, nthreads=0, out=None):
    """Exact A x in FP64. Returns y (float64); slices outside [h0,h1) stay 0."""
    lib = load()
    xin, is64 = _input(x)
    SY, SX, H = _vol_dims(xin, z_order)
    a = _angles32(angles)
    PX = max(SY, SX) if PX is None else PX
    h0, h1 = _hrange(H, h0, h1)
    y = np.zeros(y_shape(z_order, a.size, PX, H), np.float64) if out is None else out
    rc = lib.vpo_forward_f64(xin.ctypes.data, is64, y.ctypes.data, a.ctypes.data, a.size,
                             SY, SX, H, PX, z_order, h0, h1, nthreads)
    if rc:
        raise RuntimeError(f"vpo_forward_f64 failed ({rc})")
    return y


def backward_f64(y, angles, SY=None, SX=None, z_order=0, h0=0, h1=None, nthreads=0, out=None):
    """Exact AT y in FP64. Returns x (float64)."""
    lib = load()
    yin, is64 = _input(y)
    A, PX, H = _sino_dims(yin, z_order)
    a = _angles32(angles)
    if a.size != A:
        raise ValueError("len(angles) does not match the sinogram")
    SY = PX if SY is None else SY
    SX = PX if SX is None else SX
    h0, h1 = _hrange(H, h0, h1)
    x = np.zeros(x_shape(z_order, SY, SX, H), np.float64) if out is None else out
    rc = lib.vpo_backward_f64(yin.ctypes.data, is64, x.ctypes.data, a.ctypes.data, a.size,
                              SY, SX, H, PX, z_order, h0, h1, nthreads)
    if rc:
        raise RuntimeError(f"vpo_backward_f64 failed ({rc})")
    return x


def _shift64(v, n, name):
    """None, a number or n values -> None or float64 array of n values.  The
    values are rounded to float32 first, as the device library receives them."""
    if v is None:
        return None
    a = np.asarray(v, dtype=np.float32).astype(np.float64).reshape(-1)
    if a.size == 1:
        a = np.full(n, a[0])
    if a.size != n:
        raise ValueError(f"{name} needs 1 or {n} values")
    return np.ascontiguousarray(a)


def forward_geom_f64(x, angles, PX=None, z_order=0, du=None, dv=None, h0=0, h1=None, nthreads=0):
    """Exact A x in FP64 with detector translations du (pixels) / dv (slices)."""
    lib = load()
    xin, is64 = _input(x)
    SY, SX, H = _vol_dims(xin, z_order)
    a = _angles32(angles)
    PX = max(SY, SX) if PX is None else PX
    h0, h1 = _hrange(H, h0, h1)
    u, v = _shift64(du, a.size, "du"), _shift64(dv, a.size, "dv")
    y = np.zeros(y_shape(z_order, a.size, PX, H), np.float64)
    rc = lib.vpo_forward_geom_f64(xin.ctypes.data, is64, y.ctypes.data, a.ctypes.data, a.size,
                                  SY, SX, H, PX, z_order, h0, h1,
                                  None if u is None else u.ctypes.data,
                                  None if v is None else v.ctypes.data, nthreads)
    if rc:
        raise RuntimeError(f"vpo_forward_geom_f64 failed ({rc})")
    return y


def backward_geom_f64(y, angles, SY=None, SX=None, z_order=0, du=None, dv=None, h0=0, h1=None,
                      nthreads=0):
    """Exact AT y in FP64 with detector translations du / dv."""
    lib = load()
    yin, is64 = _input(y)
    A, PX, H = _sino_dims(yin, z_order)
    a = _angles32(angles)
    if a.size != A:
        raise ValueError("len(angles) does not match the sinogram")
    SY = PX if SY is None else SY
    SX = PX if SX is None else SX
    h0, h1 = _hrange(H, h0, h1)
    u, v = _shift64(du, A, "du"), _shift64(dv, A, "dv")
    x = np.zeros(x_shape(z_order, SY, SX, H), np.float64)
    rc = lib.vpo_backward_geom_f64(yin.ctypes.data, is64, x.ctypes.data, a.ctypes.data, a.size,
                                   SY, SX, H, PX, z_order, h0, h1,
                                   None if u is None else u.ctypes.data,
                                   None if v is None else v.ctypes.data, nthreads)
    if rc:
        raise RuntimeError(f"vpo_backward_geom_f64 failed ({rc})")
    return x


def ref_forward_f32(x, angles, PX=None, z_order=0, h0=0, h1=None, nthreads=0, out=None):
    """FP32 transcription of the reference forward kernel; y += A x."""
    lib = load()
    xin = np.ascontiguousarray(x, dtype=np.float32)
    SY, SX, H = _vol_dims(xin, z_order)
    a = _angles32(angles)
    PX = max(SY, SX) if PX is None else PX
    h0, h1 = _hrange(H, h0, h1)
    y = np.zeros(y_shape(z_order, a.size, PX, H), np.float32) if out is None else out
    rc = lib.vpo_ref_forward_f32(xin.ctypes.data, y.ctypes.data, a.ctypes.data, a.size,
                                 SY, SX, H, PX, z_order, h0, h1, nthreads)
    if rc:
        raise RuntimeError(f"vpo_ref_forward_f32 failed ({rc})")
    return y


def ref_backward_f32(y, angles, SY=None, SX=None, z_order=0, h0=0, h1=None, nthreads=0, out=None):
    """FP32 transcription of the reference backprojection kernel; x += AT y."""
    lib = load()
    yin = np.ascontiguousarray(y, dtype=np.float32)
    A, PX, H = _sino_dims(yin, z_order)
    a = _angles32(angles)
    if a.size != A:
        raise ValueError("len(angles) does not match the sinogram")
    SY = PX if SY is None else SY
    SX = PX if SX is None else SX
    h0, h1 = _hrange(H, h0, h1)
    x = np.zeros(x_shape(z_order, SY, SX, H), np.float32) if out is None else out
    rc = lib.vpo_ref_backward_f32(yin.ctypes.data, x.ctypes.data, a.ctypes.data, a.size,
                                  SY, SX, H, PX, z_order, h0, h1, nthreads)
    if rc:
        raise RuntimeError(f"vpo_ref_backward_f32 failed ({rc})")
    return x


def nnz(angles, SY, SX, PX, nthreads=0):
    a = _angles32(angles)
    return int(load().vpo_nnz(a.ctypes.data, a.size, SY, SX, PX, nthreads))


def weight_f64(phi, p, u):
    return float(load().vpo_weight_f64(float(phi), float(p), float(u)))


def cdf_f64(phi, t):
    return float(load().vpo_cdf_f64(float(phi), float(t)))


def angle_table_f64(angles):
    """(n, 5) table of (cos, sin, a, b, h) per angle, in FP64."""
    a = _angles32(angles)
    out = np.zeros((a.size, 5), np.float64)
    load().vpo_angle_table64(a.ctypes.data, a.size, out.ctypes.data)
    return out
