"""ctypes loader for libvoxelproj_hip.so (the HIP C-ABI library).

There is no CPU fallback: if the library is missing or cannot be loaded,
``load()`` raises, and every projector call fails loudly.
"""
import ctypes
import os
import threading

import torch  # noqa: F401  -- loads torch's HIP runtime first, so the library binds to it

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libvoxelproj_hip.so")

EXPORTS = ("vp_forward", "vp_backward", "vp_forward_f64", "vp_backward_f64",
           "vp_forward_geom", "vp_backward_geom", "vp_forward_geom_f64", "vp_backward_geom_f64",
           "vp_forward_ws", "vp_backward_ws", "vp_workspace_elems", "vp_release_workspaces",
           "vp_project_noalloc", "vp_noalloc_scratch_elems",
           "vp_vshift", "vp_forward_batched", "vp_backward_batched", "vp_transpose_batched",
           "vp_default_variant", "vp_launch_variant",
           "vp_batch_blocks", "vp_last_error", "vp_abi_version")
ABI_VERSION = 4

_lib = None
_lock = threading.Lock()


class VoxelprojLibraryError(RuntimeError):
    """The HIP projector library is missing, stale or failed to load."""


def build(verbose=False):
    """Compile the library in-tree with hipcc for gfx950."""
    from .csrc import build as _b

    global _lib
    path = _b.build(verbose=verbose)
    return path


def load():
    """Return the loaded library (loading it on first use)."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise VoxelprojLibraryError(
                f"{LIB_PATH} not found: build it with `python voxelproj_amd/csrc/build.py`"
            )
        try:
            lib = ctypes.CDLL(LIB_PATH)
        except OSError as e:  # pragma: no cover - depends on the host
            raise VoxelprojLibraryError(f"cannot load {LIB_PATH}: {e}") from e
        i64 = ctypes.c_int64
        sig = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
               i64, i64, i64, i64, ctypes.c_void_p]
        for name in ("vp_forward", "vp_backward", "vp_forward_f64", "vp_backward_f64"):
            f = getattr(lib, name)
            f.argtypes = sig
            f.restype = ctypes.c_int
        for name in ("vp_forward_geom", "vp_backward_geom", "vp_forward_geom_f64",
                     "vp_backward_geom_f64"):
            f = getattr(lib, name)
            f.argtypes = sig[:-1] + [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
                                     ctypes.c_void_p]
            f.restype = ctypes.c_int
        for name in ("vp_forward_ws", "vp_backward_ws"):
            f = getattr(lib, name)
            f.argtypes = sig[:-1] + [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
                                     ctypes.c_void_p, i64, ctypes.c_void_p]
            f.restype = ctypes.c_int
        lib.vp_workspace_elems.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, i64, i64, i64,
                                           i64, ctypes.c_int]
        lib.vp_workspace_elems.restype = i64
        lib.vp_project_noalloc.argtypes = [ctypes.c_int] + sig
        lib.vp_project_noalloc.restype = ctypes.c_int
        lib.vp_noalloc_scratch_elems.argtypes = []
        lib.vp_noalloc_scratch_elems.restype = i64
        lib.vp_release_workspaces.argtypes = []
        lib.vp_release_workspaces.restype = ctypes.c_int
        lib.vp_vshift.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                  ctypes.c_int, i64, i64, i64, i64, i64, ctypes.c_void_p,
                                  ctypes.c_int, ctypes.c_void_p]
        lib.vp_vshift.restype = ctypes.c_int
        for name in ("vp_forward_batched", "vp_backward_batched"):
            f = getattr(lib, name)
            f.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                          ctypes.c_int, ctypes.c_void_p]
            f.restype = ctypes.c_int
        lib.vp_transpose_batched.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                             ctypes.c_int, ctypes.c_void_p]
        lib.vp_transpose_batched.restype = ctypes.c_int
        lib.vp_default_variant.argtypes = [ctypes.c_int, i64]
        lib.vp_default_variant.restype = ctypes.c_int
        lib.vp_launch_variant.argtypes = [ctypes.c_int, ctypes.c_int, i64, i64, i64, i64]
        lib.vp_launch_variant.restype = ctypes.c_int
        lib.vp_batch_blocks.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, i64, i64, i64, i64]
        lib.vp_batch_blocks.restype = i64
        lib.vp_last_error.argtypes = []
        lib.vp_last_error.restype = ctypes.c_char_p
        lib.vp_abi_version.argtypes = []
        lib.vp_abi_version.restype = ctypes.c_int
        if lib.vp_abi_version() != ABI_VERSION:
            raise VoxelprojLibraryError("libvoxelproj_hip.so ABI version mismatch; rebuild it")
        _lib = lib
        return lib


def check(rc, what):
    if rc != 0:
        msg = load().vp_last_error().decode(errors="replace")
        raise RuntimeError(f"{what} failed ({rc}): {msg}")


def workspace(f64, z_order, A, SY, SX, H, PX, with_dv, device):
    """A device workspace tensor of the size one call needs (see
    vp_workspace_elems), from torch's caching allocator on the current stream."""
    import torch

    n = int(load().vp_workspace_elems(1 if f64 else 0, z_order, A, SY, SX, H, PX,
                                      1 if with_dv else 0))
    if n < 0:
        raise RuntimeError("invalid extents for a workspace")
    return torch.empty(max(n, 1), dtype=torch.float64 if f64 else torch.float32, device=device)
