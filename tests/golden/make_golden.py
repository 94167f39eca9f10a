"""Write the small seeded fixtures under tests/golden/ from the FP64 oracle.

The reference ships no golden vectors (its tests pin only adjointness at
rtol 1e-3 and container identity), so these fixtures come from this
repository's own FP64 oracle, itself pinned by tests/test_oracle.py
(polygon brute force, closed forms, invariants).  They freeze the oracle's
outputs so that a later change to it, or the GPU path, is compared with
fixed numbers.

Run:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle as vpo  # noqa: E402

CASES = [
    # name, SY, SX, H, n_angles, PX, endpoint, seed
    ("adj_full_5x6x3_a31", 6, 5, 3, 31, 11, True, 11),      # test_adjoitness.py:20-34 geometry
    ("adj_interior_5x6x3_a31", 6, 5, 3, 31, 4, True, 12),   # test_adjoitness.py:37-51 geometry
    ("identity_5x6x3_a7", 6, 5, 3, 7, 11, True, 13),        # test_identity.py:10-18 geometry
    ("grad_24x24x8_a30", 24, 24, 8, 30, 24, False, 14),     # gradcheck geometry
    ("odd_13x17x5_a45_px9", 13, 17, 5, 45, 9, False, 15),   # truncated detector, odd extents
]


def make_case(name, SY, SX, H, A, PX, endpoint, seed):
    rng = np.random.default_rng(seed)
    angles = np.linspace(0, np.pi, A, endpoint=endpoint).astype(np.float32)
    x = rng.standard_normal((SY, SX, H)).astype(np.float32)
    y = rng.standard_normal((A, PX, H)).astype(np.float32)
    Ax = vpo.forward_f64(x, angles, PX, 0)
    ATy = vpo.backward_f64(y, angles, SY, SX, 0)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), angles=angles, x=x, y=y,
                        Ax=Ax, ATy=ATy, geom=np.array([SY, SX, H, A, PX]))


def make_weight_sums():
    # examples/WoxelWeight.ipynb cells 2-3: every voxel of a 9x10 grid as a
    # unit impulse, 111 angles in [0, pi], PX = SX + SY; the per-angle sum of
    # its projection is its weight sum (exactly 1 for the exact operator).
    SX, SY, H, A = 9, 10, 1, 111
    PX = SX + SY
    angles = np.linspace(0, np.pi, A).astype(np.float32)
    sums = np.zeros((SY, SX, A))
    for iy in range(SY):
        for ix in range(SX):
            x = np.zeros((SY, SX, H), np.float32)
            x[iy, ix, 0] = 1
            sums[iy, ix] = vpo.forward_f64(x, angles, PX, 0).sum(axis=(1, 2))
    np.savez_compressed(os.path.join(HERE, "weight_sums_9x10_a111.npz"), angles=angles,
                        sums=sums, geom=np.array([SY, SX, H, A, PX]))


if __name__ == "__main__":
    for c in CASES:
        make_case(*c)
    make_weight_sums()
    print("written", sorted(f for f in os.listdir(HERE) if f.endswith(".npz")))
