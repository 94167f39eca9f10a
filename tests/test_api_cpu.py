"""Host-side logic of the Python API (no GPU needed).

Default shapes and container rules follow voxelproj/wrappers.py:10-122; the
validation that the reference lacks (documented deviations) must fire before
any device work, so all of it is testable here.
"""
import numpy as np
import pytest
import torch

import voxelproj_amd
from voxelproj_amd import _CU, wrappers


def test_default_shapes():
    assert wrappers.forward_shapes((6, 5, 3), 7, 0) == (7, 6, 3)
    assert wrappers.forward_shapes((3, 6, 5), 7, 2) == (3, 7, 6)
    assert wrappers.backward_shapes((7, 11, 3), 0) == (11, 11, 3)
    assert wrappers.backward_shapes((3, 7, 11), 2) == (3, 11, 11)


def test_bad_z_order_is_key_error_and_value_error():
    x = np.zeros((6, 5, 3), np.float32)
    for fn in (voxelproj_amd.forward, voxelproj_amd.backward):
        with pytest.raises(KeyError):
            fn(x, np.zeros(7, np.float32), None, z_order=1)
        with pytest.raises(ValueError):
            fn(x, np.zeros(7, np.float32), None, z_order=3)


def test_bad_container_type():
    x = np.zeros((6, 5, 3), np.float32)
    with pytest.raises(TypeError):
        voxelproj_amd.forward(x, np.zeros(7, np.float32), [1, 2, 3])
    with pytest.raises(TypeError):
        voxelproj_amd.backward(x, np.zeros(6, np.float32), "x")


def test_angle_count_mismatch():
    x = np.zeros((6, 5, 3), np.float32)
    with pytest.raises(ValueError, match="angle"):
        voxelproj_amd.forward(x, np.zeros(7, np.float32), (8, 11, 3))
    y = np.zeros((7, 11, 3), np.float32)
    with pytest.raises(ValueError, match="angle"):
        voxelproj_amd.backward(y, np.zeros(5, np.float32), None)


def test_h_mismatch():
    x = np.zeros((6, 5, 3), np.float32)
    with pytest.raises(ValueError, match="H mismatch"):
        voxelproj_amd.forward(x, np.zeros(7, np.float32), (7, 11, 4))
    with pytest.raises(ValueError, match="H mismatch"):
        voxelproj_amd.forward(np.zeros((3, 6, 5), np.float32), np.zeros(7, np.float32),
                              (4, 7, 11), z_order=2)


def test_non_3d_rejected():
    with pytest.raises(ValueError, match="3-D"):
        voxelproj_amd.forward(np.zeros((6, 5), np.float32), np.zeros(7, np.float32), None)


def test_cu_checks_input_like_reference():
    # projectors.cu:36-40 messages, raised as RuntimeError
    a = torch.zeros(3)
    with pytest.raises(RuntimeError, match="Tensor must be on CUDA"):
        _CU.forward_z0(torch.zeros(4, 4, 2), torch.zeros(3, 4, 2), a)
    with pytest.raises(RuntimeError, match="Tensor must be on CUDA"):
        _CU.backward_z2(np.zeros((2, 3, 4), np.float32), torch.zeros(2, 4, 4), a)


def test_zorder_error_message():
    err = wrappers.ZOrderError(5)
    assert isinstance(err, KeyError) and isinstance(err, ValueError)


def test_reference_import_names_resolve():
    # `import voxelproj` and `from voxelproj._CU import ...` (the reference's
    # voxelproj/__init__.py:1-5 and wrappers.py:4-5) resolve to this library.
    import voxelproj
    import voxelproj._CU as cu
    from voxelproj.wrappers import gpu_array_with_default_shape

    assert voxelproj.forward is voxelproj_amd.forward
    assert voxelproj.backward is voxelproj_amd.backward
    for name in ("forward_z0", "backward_z0", "forward_z2", "backward_z2"):
        assert getattr(cu, name) is getattr(_CU, name)
    assert gpu_array_with_default_shape is wrappers.gpu_array_with_default_shape
