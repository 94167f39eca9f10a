"""``voxelproj`` import name for the MI355X projector pair.

Callers of the reference package keep their imports unchanged
(voxelproj/__init__.py:1-5, voxelproj/wrappers.py:4-5):

    import voxelproj
    y = voxelproj.forward(x, angles, None)
    from voxelproj._CU import forward_z0, backward_z0, forward_z2, backward_z2

Everything here re-exports ``voxelproj_amd``; the compute path is its HIP
library for gfx950, with no CPU fallback.
"""
import torch  # noqa: F401  -- HIP runtime first, as the reference does

from voxelproj_amd import VoxelprojLibraryError, build, load  # noqa: F401

from . import _CU  # noqa: F401
from .wrappers import backward, forward  # noqa: F401

__all__ = ["forward", "backward", "build", "load", "VoxelprojLibraryError", "_CU"]
