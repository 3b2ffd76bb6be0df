"""Regression of the oracle against the committed golden fixtures
(tests/golden/*.npz, written by tests/golden/make_golden.py): the oracle must
reproduce every fixture bit for bit. The GPU side of the same fixtures is in
tests/test_gpu_golden.py."""
import glob
import os

import numpy as np
import pytest

import oracle as O

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FILES = sorted(glob.glob(os.path.join(HERE, "*.npz")))


def load(path):
    z = np.load(path)
    s = O.Scene(int(z["ni"]), int(z["nj"]), int(z["nk"]), float(z["dx"]), float(z["dt"]), z["phi"].copy(),
                z["u"].copy(), z["v"].copy(), z["w"].copy(), z["uw"].copy(), z["vw"].copy(), z["ww"].copy())
    return s, z


@pytest.mark.parametrize("path", FILES, ids=[os.path.basename(f) for f in FILES])
def test_oracle_reproduces_fixture(path):
    s, z = load(path)
    n, idx = O.assign_indices(s)
    assert n == int(z["n"])
    np.testing.assert_array_equal(idx, z["indices"])
    sys_ = O.build_matrix(s, idx, n)
    for k in ("diag", "ax", "ay", "az", "b"):
        np.testing.assert_array_equal(getattr(sys_, k), z[k], err_msg=k)
    inv_e, _ = O.ic_factor(sys_)
    np.testing.assert_array_equal(inv_e, z["inv_e"])
    res = O.solve_pressure(s, tol=1e-12, max_iter=10000)
    np.testing.assert_array_equal(res.pressure, z["pressure"])
    for f in ("u", "v", "w"):
        np.testing.assert_array_equal(getattr(res, f), z[f + "_out"])


def test_fixtures_present():
    assert len(FILES) >= 2
