"""Randomised parity: seeded random extents, angle sets (any sign, beyond
[0, pi), unsorted), both layouts, assign and accumulate, against the FP64
oracle on a few slices (the operator acts on every H slice alike, so the
first, middle and last slices stand for the rest, and they come from
different lanes, lane slots and, for H > 64, slice blocks).  The forward
kernel variants (1, 2 and 4 slices per lane) run through batched launches of
one instance on the same inputs.

Tolerances as in the parity tests: rel-L2 <= 1e-6, max-rel <= 1e-5.
"""
import numpy as np
import pytest
import torch

import oracle as vpo
from voxelproj_amd import _CU, batch

pytestmark = pytest.mark.gpu
DEV = "cuda"
REL_L2 = 1e-6
MAX_REL = 1e-5


def case(seed):
    rng = np.random.default_rng(1000 + seed)
    SY, SX = (int(v) for v in rng.integers(1, 161, 2))
    H = int(rng.integers(1, 301))
    A = int(rng.integers(1, 301))
    PX = int(rng.integers(1, 261))
    angles = rng.uniform(-2 * np.pi, 2 * np.pi, A).astype(np.float32)
    z = int(rng.integers(0, 2)) * 2
    acc = bool(rng.integers(0, 2))
    x = rng.standard_normal((SY, SX, H)).astype(np.float32)
    y = rng.standard_normal((A, PX, H)).astype(np.float32)
    return SY, SX, H, A, PX, angles, z, acc, x, y, rng


def check(got, ref, what):
    d = np.asarray(got, np.float64) - ref
    scale = np.abs(ref).max()
    r = float(np.linalg.norm(d) / max(np.linalg.norm(ref), 1e-300))
    m = float(np.abs(d).max() / max(scale, 1e-300))
    print(f"{what}: rel-L2 {r:.2e} max-rel {m:.2e}", flush=True)
    if scale == 0.0:  # nothing projects onto the detector: exact zeros
        assert np.abs(d).max() == 0.0, what
        return
    assert r <= REL_L2 and m <= MAX_REL, what


def to_z2(a):
    return np.ascontiguousarray(np.transpose(a, (2, 0, 1)))


def from_z2(a):
    return np.ascontiguousarray(np.transpose(a, (1, 2, 0)))


@pytest.mark.parametrize("seed", range(16))
def test_random_case_against_oracle(seed):
    SY, SX, H, A, PX, angles, z, acc, x, y, rng = case(seed)
    print(f"seed {seed}: SY {SY} SX {SX} H {H} A {A} PX {PX} z {z} acc {acc}", flush=True)
    hs = sorted({0, H // 2, H - 1})
    ref_y = vpo.forward_f64(np.ascontiguousarray(x[:, :, hs]), angles, PX, 0)
    ref_x = vpo.backward_f64(np.ascontiguousarray(y[:, :, hs]), angles, SY, SX, 0)
    prior_y = rng.standard_normal((A, PX, H)).astype(np.float32) if acc else np.zeros((A, PX, H), np.float32)
    prior_x = rng.standard_normal((SY, SX, H)).astype(np.float32) if acc else np.zeros((SY, SX, H), np.float32)
    at = torch.tensor(angles, device=DEV)
    if z == 0:
        yo = torch.tensor(prior_y, device=DEV)
        _CU.forward_z0(torch.tensor(x, device=DEV), yo, at)
        xo = torch.tensor(prior_x, device=DEV)
        _CU.backward_z0(torch.tensor(y, device=DEV), xo, at)
        got_y, got_x = yo.cpu().numpy(), xo.cpu().numpy()
    else:
        yo = torch.tensor(to_z2(prior_y), device=DEV)
        _CU.forward_z2(torch.tensor(to_z2(x), device=DEV), yo, at)
        xo = torch.tensor(to_z2(prior_x), device=DEV)
        _CU.backward_z2(torch.tensor(to_z2(y), device=DEV), xo, at)
        got_y, got_x = from_z2(yo.cpu().numpy()), from_z2(xo.cpu().numpy())
    # subtracting the prior is exact enough here: |prior| ~ |op| ~ 1..100
    check(got_y[:, :, hs] - prior_y[:, :, hs], ref_y, f"A seed {seed}")
    check(got_x[:, :, hs] - prior_x[:, :, hs], ref_x, f"AT seed {seed}")
    xt = torch.tensor(x, device=DEV)
    for v in batch.FORWARD_VARIANTS:
        yv = torch.empty((A, PX, H), device=DEV)
        batch.BatchPlan("A", [xt], [yv], [at], accumulate=False, variant=v).launch()
        check(yv.cpu().numpy()[:, :, hs], ref_y, f"A variant {v} seed {seed}")
