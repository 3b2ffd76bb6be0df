"""The GPU against the committed golden fixtures (tests/golden/*.npz, written
by the oracle through tests/golden/make_golden.py), through the C ABI:

- projection fixtures: indices, A, b and the IC(0) pivots bit for bit
  (elq_assign_indices, elq_build_matrix, elq_ic_factor), the pressure of elq_project at rel.res 1e-10 within
  2e-6 of the fixture's (FP64; mixed 2e-5) relative to max|p|, and the
  updated velocities within that bound times 2 dt / (0.01 dx), the largest
  factor the update applies to a pressure difference;
- the frame fixture, bit for bit: face flags, extrapolate_velocity,
  advect_level_set, advect_velocity and the reinitialisation (with its pass
  count).
"""
import glob
import os

import numpy as np
import pytest

import elq

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FILES = sorted(f for f in glob.glob(os.path.join(HERE, "*.npz")) if not os.path.basename(f).startswith("frame"))
FRAME = os.path.join(HERE, "frame16.npz")


def scene(z):
    return elq.Scene(int(z["ni"]), int(z["nj"]), int(z["nk"]), float(z["dx"]), float(z["dt"]), z["phi"].copy(),
                     z["u"].copy(), z["v"].copy(), z["w"].copy(), z["uw"].copy(), z["vw"].copy(), z["ww"].copy())


def same(a, b):
    return np.array_equal(a.view(np.uint32), b.view(np.uint32)) if a.dtype == np.float32 else np.array_equal(a, b)


@pytest.mark.parametrize("path", FILES, ids=[os.path.basename(f) for f in FILES])
def test_projection_fixture(ctx, path):
    z = np.load(path)
    s = scene(z)
    n, idx = ctx.assign_indices(s)
    assert n == int(z["n"])
    np.testing.assert_array_equal(idx, z["indices"])
    m = ctx.build_matrix(s, n)
    for k in ("diag", "ax", "ay", "az", "b"):
        np.testing.assert_array_equal(m[k].view(np.uint64), z[k].view(np.uint64), err_msg=k)
    # IC(0) pivots of the preconditioner set-up (simulator.cpp:102-107), bit for bit
    inv_e = ctx.ic_factor(s, n)
    np.testing.assert_array_equal(inv_e.view(np.uint64), z["inv_e"].view(np.uint64))
    scale = np.abs(z["pressure"]).max()
    for prec, tol in ((elq.FP64, 2e-6), (elq.MIXED, 2e-5)):
        t = scene(z)
        p, _, st = ctx.solve_pressure(t, elq.default_opts(tol=1e-10, max_iter=10000, precision=prec))
        assert st.converged
        assert np.abs(p - z["pressure"]).max() <= tol * scale
        for k in ("u", "v", "w"):
            # u -= dt (p_r - p_l) / (dx theta), theta >= 0.01 (projection.cpp:93-134)
            assert np.abs(getattr(t, k) - z[k + "_out"]).max() <= tol * scale * s.dt / s.dx * 200, k


def test_frame_fixture(ctx):
    z = np.load(FRAME)
    ni, nj, nk = int(z["ni"]), int(z["nj"]), int(z["nk"])
    dx, dt = np.float32(z["dx"]), np.float32(z["dt"])
    L = ctx.L
    host = {k: z[k].copy() for k in ("phi", "u", "v", "w", "uw", "vw", "ww")}
    dev = {k: ctx.alloc(a.nbytes) for k, a in host.items()}
    flags = {k: ctx.alloc(z[k + "_flags"].size) for k in "uvw"}
    try:
        for k, a in host.items():
            ctx.h2d(dev[k], a)
        g = elq.Grid()
        g.ni, g.nj, g.nk, g.dx = ni, nj, nk, dx
        g.level_set, g.u, g.v, g.w = dev["phi"], dev["u"], dev["v"], dev["w"]
        g.u_weight, g.v_weight, g.w_weight = dev["uw"], dev["vw"], dev["ww"]

        def get(p, like):
            out = np.empty_like(like)
            ctx.d2h(out, p)
            return out

        assert L.elq_find_valid_velocity_device(ctx.h, g, flags["u"], flags["v"], flags["w"]) == elq.OK
        for k in "uvw":
            assert same(get(flags[k], z[k + "_flags"]), z[k + "_flags"]), k
        assert L.elq_extrapolate_velocity_device(ctx.h, g, flags["u"], flags["v"], flags["w"], 10) == elq.OK
        for k in "uvw":
            assert same(get(dev[k], z[k + "_ext"]), z[k + "_ext"]), k
            assert same(get(flags[k], z[k + "_ext_flags"]), z[k + "_ext_flags"]), k
        assert L.elq_advect_level_set_device(ctx.h, ni, nj, nk, dx, dt, dev["phi"], dev["u"], dev["v"],
                                             dev["w"]) == elq.OK
        assert L.elq_advect_velocity_device(ctx.h, g, dt) == elq.OK
        for k in ("phi", "u", "v", "w"):
            assert same(get(dev[k], z[k + "_adv"]), z[k + "_adv"]), k
    finally:
        for p in list(dev.values()) + list(flags.values()):
            ctx.free(p)
    rei = z["phi_adv"].copy()
    left, passes = ctx.reinitialize(rei, z["ls_flags"].copy(), (ni, nj, nk), float(dx))
    assert (left, passes) == (int(z["rei_unreached"]), int(z["rei_passes"]))
    assert same(rei, z["phi_rei"])
