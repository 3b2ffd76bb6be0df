"""Level-set reinitialisation on the GPU (reinitialization.cpp:19-156),
through the C ABI, against the oracle, bit for bit (float32 words), with the
same count of passes and of unreached cells:

- planar fronts along an axis, at 45 degrees and along a body diagonal (the
  1-, 2- and 3-neighbour upwind formulas and the geometric interface
  distance);
- meeting fronts (the -x/-y/-z neighbour is taken first);
- a sphere with noise on a room() grid, flags from find_valid_level_set, and
  a pocket of cells walled off by invalid cells;
- a wall of invalid cells with a hole, so that cells behind it are reached
  by a detour (longer than their taxicab distance), also with the tile
  route's stamp rebase forced every 1 to 80 passes;
- the 128^3 drop scene's level set after one advection step (about 100
  passes, several host batches);
- grids one or two cells thick on some axis, and fronts along an axis of
  4092 cells.

The library takes one of two routes, both checked here: passes ordered by
the taxicab distance to the interface where that distance is the pass count
(every test but the pocket and the wall), tile passes otherwise.
"""
import numpy as np
import pytest

import elq
import oracle as O

pytestmark = pytest.mark.gpu


def same(a, b):
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


def check(ctx, phi, valid, dims, dx, route="rei_layer"):
    ref = phi.copy()
    r_left, r_passes = O.reinitialize(ref, valid, dims, dx)
    got = phi.copy()
    ctx.set_profiling(True)
    try:
        left, passes = ctx.reinitialize(got, valid, dims, dx)
        kinds = ctx.kernel_times()
    finally:
        ctx.set_profiling(False)
    assert (left, passes) == (r_left, r_passes)
    if r_passes:
        assert route in kinds, sorted(kinds)
    assert same(got, ref), np.flatnonzero(got.view(np.uint32) != ref.view(np.uint32))[:10]
    return passes


def xyz(dims, dx):
    ni, nj, nk = dims
    k, j, i = np.meshgrid(np.arange(nk), np.arange(nj), np.arange(ni), indexing="ij")
    return ((i + 0.5) * dx).reshape(-1), ((j + 0.5) * dx).reshape(-1), ((k + 0.5) * dx).reshape(-1)


@pytest.mark.parametrize("normal,scale", [((0, 1, 0), 3.0), ((1, 1, 0), 2.0), ((1, 1, 1), 1.7), ((-1, 2, 1), 0.6)])
def test_planar_fronts_bit_exact(ctx, normal, scale):
    dims, dx = (40, 36, 44), 1.0 / 40
    x, y, z = xyz(dims, dx)
    nv = np.array(normal, float) / np.linalg.norm(normal)
    c = nv @ np.array([0.5, 0.45, 0.55]) + 0.3 * dx
    phi = (scale * (nv[0] * x + nv[1] * y + nv[2] * z - c)).astype(np.float32)
    assert check(ctx, phi, np.ones(phi.size, np.uint8), dims, dx) > 10


def test_meeting_fronts_bit_exact(ctx):
    dims, dx = (33, 30, 35), 1.0 / 32
    x, y, z = xyz(dims, dx)
    phi = (2.0 * np.minimum(np.minimum(x - 0.31, 0.72 - x), np.minimum(y - 0.2, 0.83 - z))).astype(np.float32)
    check(ctx, phi, np.ones(phi.size, np.uint8), dims, dx)


def test_sphere_with_noise_on_room_and_pocket(ctx):
    dims, dx = (48, 40, 44), 1.0 / 48
    ni, nj, nk = dims
    x, y, z = xyz(dims, dx)
    rng = np.random.default_rng(3)
    phi = (np.sqrt((x - 0.5) ** 2 + (y - 0.4) ** 2 + (z - 0.45) ** 2) - 0.25).astype(np.float32)
    phi *= rng.uniform(0.5, 1.5, phi.size).astype(np.float32)
    k, j, i = np.meshgrid(np.arange(nk), np.arange(nj), np.arange(ni), indexing="ij")
    shell = (np.abs(i - 40) <= 2) & (np.abs(j - 33) <= 2) & (np.abs(k - 36) <= 2)
    pocket = (np.abs(i - 40) <= 1) & (np.abs(j - 33) <= 1) & (np.abs(k - 36) <= 1)
    phi[(rng.random(phi.size) < 0.01) & ~shell.reshape(-1)] = 0.0   # no interface inside the pocket
    nodal = np.empty((ni + 1) * (nj + 1) * (nk + 1), np.float32)
    O.lib().orc_room_nodal(ni, nj, nk, nodal)
    valid = np.empty(phi.size, np.uint8)
    O.lib().orc_find_valid_level_set(ni, nj, nk, nodal, valid)
    # a walled-off pocket of air cells
    valid[(shell & ~pocket).reshape(-1)] = 0
    ref = phi.copy()
    left, _ = O.reinitialize(ref, valid, dims, dx)
    assert left == 27
    check(ctx, phi, valid, dims, dx, route="rei_pass")


@pytest.mark.parametrize("dims", [(1, 30, 25), (27, 1, 19), (2, 3, 70), (65, 2, 1)])
def test_thin_grids_bit_exact(ctx, dims):
    """Grids one or two cells thick on some axis: the neighbour masks of the
    distance route never point across those faces."""
    dx = 1.0 / 30
    x, y, z = xyz(dims, dx)
    rng = np.random.default_rng(sum(dims))
    phi = ((np.sin(7 * x + 3 * y) + np.cos(5 * z) - 0.4) * rng.uniform(0.5, 1.5, x.size)).astype(np.float32)
    assert check(ctx, phi, np.ones(phi.size, np.uint8), dims, dx) >= 10


@pytest.mark.parametrize("flags", ["all_valid", "none_valid"])
def test_no_interface(ctx, flags):
    """No sign change anywhere: every valid cell is unreached (the reference
    would loop forever), no pass runs; with no valid cell nothing happens."""
    dims, dx = (20, 18, 22), 1.0 / 20
    n = int(np.prod(dims))
    phi = np.full(n, 0.3, np.float32)
    valid = np.ones(n, np.uint8) if flags == "all_valid" else np.zeros(n, np.uint8)
    got = phi.copy()
    left, passes = ctx.reinitialize(got, valid, dims, dx)
    ref = phi.copy()
    assert (left, passes) == O.reinitialize(ref, valid, dims, dx) == ((n if flags == "all_valid" else 0), 0)
    assert same(got, ref)


def test_wall_with_hole_detour_bit_exact(ctx):
    dims, dx = (36, 32, 40), 1.0 / 36
    ni, nj, nk = dims
    x, y, z = xyz(dims, dx)
    phi = (1.3 * (y - 0.3)).astype(np.float32)   # liquid below y = 0.3
    k, j, i = np.meshgrid(np.arange(nk), np.arange(nj), np.arange(ni), indexing="ij")
    wall = (j == 16) & ~((np.abs(i - 6) <= 1) & (np.abs(k - 30) <= 1))
    valid = np.ones(phi.size, np.uint8)
    valid[wall.reshape(-1)] = 0
    ref = phi.copy()
    left, passes = O.reinitialize(ref, valid, dims, dx)
    assert left == 0 and passes > 40   # 21 passes reach the far corner without the wall
    check(ctx, phi, valid, dims, dx, route="rei_pass")


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_front_along_a_4092_cell_axis(ctx, axis):
    # a plane across the middle of a 4092-cell axis: about 2046 passes, the
    # distance route's lists and 16-bit distances near their longest
    dims = [6, 6, 6]
    dims[axis] = 4092
    dx = 1.0 / 4092
    x, y, z = xyz(tuple(dims), dx)
    coord = (x, y, z)[axis]
    phi = (1.7 * (coord - 0.5 - 0.3 * dx)).astype(np.float32)
    valid = np.ones(phi.size, np.uint8)
    assert check(ctx, phi, valid, tuple(dims), dx) > 2000


def wall_with_hole():
    dims, dx = (36, 32, 40), 1.0 / 36
    ni, nj, nk = dims
    x, y, z = xyz(dims, dx)
    phi = (1.3 * (y - 0.3)).astype(np.float32)   # liquid below y = 0.3
    k, j, i = np.meshgrid(np.arange(nk), np.arange(nj), np.arange(ni), indexing="ij")
    wall = (j == 16) & ~((np.abs(i - 6) <= 1) & (np.abs(k - 30) <= 1))
    valid = np.ones(phi.size, np.uint8)
    valid[wall.reshape(-1)] = 0
    return phi, valid, dims, dx


@pytest.mark.parametrize("every", [1, 2, 5, 38, 76, 77, 78, 80])
def test_tile_route_stamp_rebase_bit_exact(ctx, every):
    # The tile route folds its 16-bit stamps back to 0 every `every` passes
    # (60000 by default, more than any test reaches). The wall scene's march
    # takes 77 passes in host batches of 32: rebases before, at and after
    # its last pass (also inside the batch that runs on past the end) must
    # leave values, pass count and unreached count unchanged.
    phi, valid, dims, dx = wall_with_hole()
    ctx.set_reinit_rebase(every)
    try:
        assert check(ctx, phi, valid, dims, dx, route="rei_pass") == 77
    finally:
        ctx.set_reinit_rebase(60000)


def test_reinit_rebase_interval_range(ctx):
    for bad in (0, -1, 60001):
        with pytest.raises(elq.ElqError) as e:
            ctx.set_reinit_rebase(bad)
        assert e.value.code == elq.INVALID_ARG


def test_drop128_after_advection_bit_exact(ctx):
    s = elq.scene_drop(128)
    ctx.solve_pressure(s, elq.default_opts(tol=1e-6, max_iter=5000, precision=elq.MIXED))
    phi = np.empty_like(s.phi)
    O.lib().orc_advect_level_set(s.ni, s.nj, s.nk, np.float32(s.dx), np.float32(s.dt), s.phi, s.u, s.v, s.w, phi)
    ni, nj, nk = s.ni, s.nj, s.nk
    nodal = np.empty((ni + 1) * (nj + 1) * (nk + 1), np.float32)
    O.lib().orc_room_nodal(ni, nj, nk, nodal)
    valid = np.empty(phi.size, np.uint8)
    O.lib().orc_find_valid_level_set(ni, nj, nk, nodal, valid)
    assert check(ctx, phi, valid, (ni, nj, nk), s.dx) > 40
