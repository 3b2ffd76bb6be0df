"""Regression of the oracle against the committed golden fixtures
(tests/golden/*.npz, written by tests/golden/make_golden.py): the oracle must
reproduce every fixture bit for bit, the projection fixtures and the frame
fixture (extrapolation, advection, reinitialisation after a projection). The
GPU side of the same fixtures is in tests/test_gpu_golden.py."""
import glob
import os

import numpy as np
import pytest

import oracle as O

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FILES = sorted(f for f in glob.glob(os.path.join(HERE, "*.npz")) if not os.path.basename(f).startswith("frame"))
FRAME = os.path.join(HERE, "frame16.npz")


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
    assert len(FILES) >= 2 and os.path.exists(FRAME)


def test_oracle_reproduces_frame_fixture():
    z = np.load(FRAME)
    ni, nj, nk, dx, dt = int(z["ni"]), int(z["nj"]), int(z["nk"]), float(z["dx"]), float(z["dt"])
    s = O.Scene(ni, nj, nk, dx, dt, z["phi"].copy(), z["u"].copy(), z["v"].copy(), z["w"].copy(), z["uw"].copy(),
                z["vw"].copy(), z["ww"].copy())
    flags = O.find_valid_velocity(s)
    for k, f in zip("uvw", flags):
        np.testing.assert_array_equal(f, z[k + "_flags"])
    dims = ((ni + 1, nj, nk), (ni, nj + 1, nk), (ni, nj, nk + 1))
    ext = [z[k].copy() for k in ("u", "v", "w")]
    for k, a, f, d in zip("uvw", ext, flags, dims):
        f = f.copy()
        O.extrapolate(a, f, d, 10)
        np.testing.assert_array_equal(a.view(np.uint32), z[k + "_ext"].view(np.uint32))
        np.testing.assert_array_equal(f, z[k + "_ext_flags"])
    adv = O.Scene(ni, nj, nk, dx, dt, z["phi"].copy(), *[a.copy() for a in ext], z["uw"], z["vw"], z["ww"])
    O.advect_level_set(adv, dt)
    O.advect_velocity(adv, dt)
    for k in ("phi", "u", "v", "w"):
        np.testing.assert_array_equal(getattr(adv, k).view(np.uint32), z[k + "_adv"].view(np.uint32), err_msg=k)
    rei = z["phi_adv"].copy()
    left, passes = O.reinitialize(rei, z["ls_flags"].copy(), (ni, nj, nk), dx)
    assert (left, passes) == (int(z["rei_unreached"]), int(z["rei_passes"]))
    np.testing.assert_array_equal(rei.view(np.uint32), z["phi_rei"].view(np.uint32))
