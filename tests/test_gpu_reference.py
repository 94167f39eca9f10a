"""The HIP projectors against the reference's own algorithm.

The reference CUDA operator cannot run here; its FP32 arithmetic is restated
by the oracle's transcription (oracle.ref_forward_f32 / ref_backward_f32,
following sp_forward_z0.cuh:8-185 and sp_backward_z0.cuh:8-100 and their
z_order 2 twins).  Two checks:

* on the reference's own test geometries (test/test_adjoitness.py:12-51,
  test/test_identity.py:10-18) the HIP results equal the transcription's
  within rel-L2 1e-5, in both layouts and both operators;
* at the benchmark sizes the transcription drifts from the exact operator
  (tests/test_reference_drift.py), and the HIP results are closer to the
  exact FP64 operator than the reference algorithm is.
"""
import numpy as np
import pytest
import torch

import oracle as vpo
from voxelproj_amd import _CU

pytestmark = pytest.mark.gpu
DEV = "cuda"
REF_REL_L2 = 1e-5


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def _gpu(op, inp, angles, out_shape, z):
    t = torch.as_tensor(np.ascontiguousarray(inp), device=DEV)
    o = torch.zeros(out_shape, device=DEV)
    a = torch.as_tensor(np.asarray(angles, np.float32), device=DEV)
    getattr(_CU, f"{op}_z{z}")(t, o, a)
    return o.cpu().numpy()


# (SY, SX, H, n_angles, PX): the adjointness cases (full detector PX = SX + SY
# and truncated PX = min - 1) and the identity case
REFERENCE_GEOMETRIES = [(6, 5, 3, 31, 11), (6, 5, 3, 31, 4), (6, 5, 3, 7, 11)]


@pytest.mark.parametrize("z", [0, 2])
@pytest.mark.parametrize("geom", REFERENCE_GEOMETRIES)
def test_matches_reference_algorithm_on_its_test_geometries(geom, z):
    SY, SX, H, A, PX = geom
    ang = np.linspace(0, np.pi, A).astype(np.float32)
    rng = np.random.default_rng(128437)
    xs = (SY, SX, H) if z == 0 else (H, SY, SX)
    ys = (A, PX, H) if z == 0 else (H, A, PX)
    for data in ("uniform", "normal"):
        x = (rng.random(xs) if data == "uniform" else rng.standard_normal(xs)).astype(np.float32)
        y = (rng.random(ys) if data == "uniform" else rng.standard_normal(ys)).astype(np.float32)
        ea = rel_l2(_gpu("forward", x, ang, ys, z), vpo.ref_forward_f32(x, ang, PX, z))
        eb = rel_l2(_gpu("backward", y, ang, xs, z), vpo.ref_backward_f32(y, ang, SY, SX, z))
        print(f"{geom} z{z} {data}: vs reference algorithm A {ea:.2e}, AT {eb:.2e}", flush=True)
        assert ea <= REF_REL_L2 and eb <= REF_REL_L2


@pytest.mark.parametrize("N,H,h1", [(128, 128, 128), (512, 512, 2)])
def test_closer_to_exact_than_reference_algorithm(N, H, h1):
    # h1 < H: the CPU legs cover slices [0, h1) only (slices are independent)
    A = 192
    ang = np.linspace(0, np.pi, A).astype(np.float32)
    g = torch.Generator().manual_seed(7)
    x = torch.randn((N, N, H), generator=g).numpy()
    y = torch.randn((A, N, H), generator=g).numpy()
    ya = _gpu("forward", x, ang, (A, N, H), 0)[:, :, :h1]
    xb = _gpu("backward", y, ang, (N, N, H), 0)[:, :, :h1]
    ex_a = vpo.forward_f64(x, ang, N, 0, 0, h1)[:, :, :h1]
    ex_b = vpo.backward_f64(y, ang, N, N, 0, 0, h1)[:, :, :h1]
    ref_a = vpo.ref_forward_f32(x, ang, N, 0, 0, h1)[:, :, :h1]
    ref_b = vpo.ref_backward_f32(y, ang, N, N, 0, 0, h1)[:, :, :h1]
    hip_a, hip_b = rel_l2(ya, ex_a), rel_l2(xb, ex_b)
    drift_a, drift_b = rel_l2(ref_a, ex_a), rel_l2(ref_b, ex_b)
    print(f"{N}^2 x {H} (slices < {h1}) x {A}: vs exact, HIP A {hip_a:.2e} AT {hip_b:.2e}; "
          f"reference algorithm A {drift_a:.2e} AT {drift_b:.2e}; HIP vs reference algorithm "
          f"A {rel_l2(ya, ref_a):.2e} AT {rel_l2(xb, ref_b):.2e}", flush=True)
    assert hip_a <= 1e-6 and hip_b <= 1e-6
    assert hip_a < drift_a and hip_b < drift_b
