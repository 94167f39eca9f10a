"""``voxelproj._CU``: the compiled-module entry points
(voxelproj/csrc/voxelproj.cpp:11-34), served by ``voxelproj_amd._CU``."""
from voxelproj_amd._CU import backward_z0, backward_z2, forward_z0, forward_z2  # noqa: F401

__all__ = ["forward_z0", "backward_z0", "forward_z2", "backward_z2"]
