"""``voxelproj.wrappers`` (voxelproj/wrappers.py:10-122), served by
``voxelproj_amd.wrappers``."""
from voxelproj_amd.wrappers import backward, forward, gpu_array_with_default_shape  # noqa: F401

__all__ = ["forward", "backward", "gpu_array_with_default_shape"]
