"""MI355X-native exact-trapezoid parallel-beam projector pair (A / AT).

Drop-in for the reference package ``voxelproj`` (voxelproj/__init__.py:1-5):

    import voxelproj_amd as voxelproj
    y = voxelproj.forward(x, angles, y=None, z_order=0)
    x = voxelproj.backward(y, angles, x=None, z_order=0)

The compute path is the hand-written HIP library ``libvoxelproj_hip.so``
(gfx950) behind a C ABI; there is no CPU fallback.
"""
import torch  # noqa: F401  -- HIP runtime first

from . import _CU  # noqa: F401
from ._lib import VoxelprojLibraryError, build, load  # noqa: F401
from .wrappers import backward, forward  # noqa: F401

__all__ = ["forward", "backward", "build", "load", "VoxelprojLibraryError", "_CU"]
