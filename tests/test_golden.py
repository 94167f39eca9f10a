"""The oracle reproduces the committed fixtures (tests/golden/*.npz)."""
import glob
import os

import numpy as np
import pytest

import oracle as vpo

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = sorted(p for p in glob.glob(os.path.join(HERE, "*.npz")) if "weight_sums" not in p)


def test_fixtures_present():
    assert len(CASES) >= 5
    assert os.path.exists(os.path.join(HERE, "weight_sums_9x10_a111.npz"))


@pytest.mark.parametrize("path", CASES, ids=[os.path.basename(p) for p in CASES])
def test_oracle_reproduces_fixture(path):
    d = np.load(path)
    SY, SX, H, A, PX = (int(v) for v in d["geom"])
    np.testing.assert_allclose(vpo.forward_f64(d["x"], d["angles"], PX, 0), d["Ax"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(vpo.backward_f64(d["y"], d["angles"], SY, SX, 0), d["ATy"], rtol=0, atol=1e-12)


def test_weight_sums_fixture_is_unity():
    d = np.load(os.path.join(HERE, "weight_sums_9x10_a111.npz"))
    np.testing.assert_allclose(d["sums"], 1.0, atol=1e-12)
