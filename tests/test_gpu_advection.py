"""Semi-Lagrangian advection on the GPU (advection.cpp:9-27), through the C
ABI, against the oracle, bit for bit (float32 words):

- elq_advect_velocity_device / elq_advect_level_set_device in the
  simulator's form (simulator.cpp:79-87) on seeded scenes with random face
  velocities, a non-dyadic dx (1/40), an anisotropic grid that is no tile
  multiple, a step long enough that backtraces leave the grid (clamped
  indices), and grids with an axis of one cell;
- elq_advect_staggered_device / elq_advect_collocated_device with separate
  carried values and carrying velocity;
- the 128^3 drop scene after one projection, both advections;
- outputs aliasing inputs are refused.
"""
import numpy as np
import pytest

import elq
import oracle as O

pytestmark = pytest.mark.gpu


def same(a, b):
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


class Dev:
    def __init__(self, ctx, **arrays):
        self.ctx, self.p, self.h = ctx, {}, arrays
        for k, a in arrays.items():
            self.p[k] = ctx.alloc(max(a.nbytes, 4))
            ctx.h2d(self.p[k], a)

    def get(self, k):
        out = np.empty_like(self.h[k])
        self.ctx.d2h(out, self.p[k])
        return out

    def close(self):
        for p in self.p.values():
            self.ctx.free(p)


def random_fields(ni, nj, nk, seed, amp):
    rng = np.random.default_rng(seed)
    f = lambda n: rng.uniform(-amp, amp, n).astype(np.float32)
    return f((ni + 1) * nj * nk), f(ni * (nj + 1) * nk), f(ni * nj * (nk + 1)), rng


CASES = [((40, 40, 40), 1 / 40, 0.01, 1.0, 1), ((23, 17, 29), 1 / 29, 0.02, 0.8, 2),
         ((16, 16, 16), 1 / 16, 0.5, 1.0, 3),    # leaves the grid
         # an axis of one cell (every index on it clamps to 0), large non-dyadic steps
         ((1, 7, 5), 1 / 7, 0.05, 1.0, 4), ((33, 1, 2), 0.37, 0.9, 2.0, 5), ((3, 2, 1), 0.3, 0.2, 1.5, 6),
         # axes of 4092 cells
         ((4092, 3, 2), 1 / 4092, 0.01, 1.0, 7), ((2, 4092, 3), 1 / 64, 0.05, 1.0, 8),
         ((3, 2, 4092), 1 / 64, 0.05, 1.0, 9)]


@pytest.mark.parametrize("dims,dx,dt,amp,seed", CASES)
def test_simulator_advections_bit_exact(ctx, dims, dx, dt, amp, seed):
    ni, nj, nk = dims
    u, v, w, rng = random_fields(ni, nj, nk, seed, amp)
    phi = rng.uniform(-1, 1, ni * nj * nk).astype(np.float32)
    # oracle: advect_level_set, then advect_velocity (apps/eulerian_liquid_3d.cpp:111-112)
    rphi = np.empty_like(phi)
    O.lib().orc_advect_level_set(ni, nj, nk, np.float32(dx), np.float32(dt), phi, u, v, w, rphi)
    ru, rv, rw = np.empty_like(u), np.empty_like(v), np.empty_like(w)
    O.lib().orc_advect_velocity(ni, nj, nk, np.float32(dx), np.float32(dt), u, v, w, ru, rv, rw)
    d = Dev(ctx, phi=phi, u=u, v=v, w=w)
    try:
        L = ctx.L
        assert L.elq_advect_level_set_device(ctx.h, ni, nj, nk, np.float32(dx), np.float32(dt), d.p["phi"],
                                             d.p["u"], d.p["v"], d.p["w"]) == elq.OK
        assert same(d.get("phi"), rphi)
        g = elq.Grid()
        g.ni, g.nj, g.nk, g.dx = ni, nj, nk, np.float32(dx)
        g.u, g.v, g.w = d.p["u"], d.p["v"], d.p["w"]
        assert L.elq_advect_velocity_device(ctx.h, g, np.float32(dt)) == elq.OK
        for k, r in zip("uvw", (ru, rv, rw)):
            assert same(d.get(k), r), k
        assert not same(rphi, phi)
    finally:
        d.close()


def test_general_forms_bit_exact(ctx):
    ni, nj, nk, dx, dt = 24, 20, 18, np.float32(1 / 24), np.float32(0.03)
    vel = random_fields(ni, nj, nk, 7, 1.0)[:3]
    val = random_fields(ni, nj, nk, 8, 5.0)[:3]
    phi = np.random.default_rng(9).uniform(-1, 1, ni * nj * nk).astype(np.float32)
    ref = [np.empty_like(a) for a in val]
    O.lib().orc_advect_staggered(ni, nj, nk, dx, dt, *vel, *val, *ref)
    rphi = np.empty_like(phi)
    O.lib().orc_advect_level_set(ni, nj, nk, dx, dt, phi, *vel, rphi)
    d = Dev(ctx, vu=vel[0], vv=vel[1], vw=vel[2], au=val[0], av=val[1], aw=val[2], phi=phi,
            ou=np.zeros_like(val[0]), ov=np.zeros_like(val[1]), ow=np.zeros_like(val[2]), ophi=np.zeros_like(phi))
    try:
        p, L = d.p, ctx.L
        assert L.elq_advect_staggered_device(ctx.h, ni, nj, nk, dx, dt, p["ou"], p["ov"], p["ow"], p["au"], p["av"],
                                             p["aw"], p["vu"], p["vv"], p["vw"]) == elq.OK
        for k, r in zip(("ou", "ov", "ow"), ref):
            assert same(d.get(k), r), k
        assert L.elq_advect_collocated_device(ctx.h, ni, nj, nk, dx, dt, p["ophi"], p["phi"], p["vu"], p["vv"],
                                              p["vw"]) == elq.OK
        assert same(d.get("ophi"), rphi)
        # aliasing outputs and inputs is refused
        assert L.elq_advect_collocated_device(ctx.h, ni, nj, nk, dx, dt, p["phi"], p["phi"], p["vu"], p["vv"],
                                              p["vw"]) == elq.INVALID_ARG
        assert L.elq_advect_staggered_device(ctx.h, ni, nj, nk, dx, dt, p["au"], p["ov"], p["ow"], p["au"], p["av"],
                                             p["aw"], p["vu"], p["vv"], p["vw"]) == elq.INVALID_ARG
    finally:
        d.close()


def test_drop128_after_projection_bit_exact(ctx):
    s = elq.scene_drop(128)
    p, idx, st = ctx.solve_pressure(s, elq.default_opts(tol=1e-6, max_iter=5000, precision=elq.MIXED))
    ni, nj, nk, dx, dt = s.ni, s.nj, s.nk, np.float32(s.dx), np.float32(s.dt)
    rphi = np.empty_like(s.phi)
    O.lib().orc_advect_level_set(ni, nj, nk, dx, dt, s.phi, s.u, s.v, s.w, rphi)
    ru, rv, rw = np.empty_like(s.u), np.empty_like(s.v), np.empty_like(s.w)
    O.lib().orc_advect_velocity(ni, nj, nk, dx, dt, s.u, s.v, s.w, ru, rv, rw)
    d = Dev(ctx, phi=s.phi.copy(), u=s.u.copy(), v=s.v.copy(), w=s.w.copy())
    try:
        L = ctx.L
        assert L.elq_advect_level_set_device(ctx.h, ni, nj, nk, dx, dt, d.p["phi"], d.p["u"], d.p["v"],
                                             d.p["w"]) == elq.OK
        g = elq.Grid()
        g.ni, g.nj, g.nk, g.dx = ni, nj, nk, dx
        g.u, g.v, g.w = d.p["u"], d.p["v"], d.p["w"]
        assert L.elq_advect_velocity_device(ctx.h, g, dt) == elq.OK
        assert same(d.get("phi"), rphi)
        for k, r in zip("uvw", (ru, rv, rw)):
            assert same(d.get(k), r), k
    finally:
        d.close()
