"""Pins of the oracle's semi-Lagrangian advection (advection.cpp:9-27,
ode.h:30-35, interpolation.h:10-72,186-193) against facts fixed without the
oracle's formula:

- zero velocity leaves every grid bit-identical when dx is a power of two
  (with dx = 1/12 the reference's own (pos - origin - dx/2) / dx misses the
  integer by an ulp, and the interpolation then mixes neighbours);
- a uniform velocity of one cell per step (u = 1, dt = dx = 1/16) shifts the
  level set by exactly one cell downstream on each axis, with the upstream
  boundary clamped (grid.h:370-379); three cells per step clamps twice as far.
  A wrong sign of the backtrace, a swapped axis or a half-cell origin error
  breaks the exact shift;
- a linear velocity field (solid rotation about the z axis) and a linear
  level set: trilinear interpolation reproduces linear functions, so the
  result is the closed-form RK2 midpoint backtrace
  x_b = x - dt A r + dt^2/2 A^2 r (r = x - c), to float rounding. RK1 (no
  dt^2 term), a midpoint factor other than 1/2 or a half-cell shift of any
  grid's origin miss it by 1e-5 or more; the velocity grids advected by
  themselves follow the same closed form at their own face positions.
"""
import numpy as np
import pytest

import oracle as O


def scene(n, dx):
    s = O.make_scene(n, n, n, 0.4, [])
    s.dx = dx
    return s


def face_positions(axis, n, dx):
    dims = [n, n, n]
    dims[axis] += 1
    k, j, i = np.meshgrid(np.arange(dims[2]), np.arange(dims[1]), np.arange(dims[0]), indexing="ij")
    x = (i + 0.5) * dx - (0.5 * dx if axis == 0 else 0.0)
    y = (j + 0.5) * dx - (0.5 * dx if axis == 1 else 0.0)
    z = (k + 0.5) * dx - (0.5 * dx if axis == 2 else 0.0)
    return x.reshape(-1), y.reshape(-1), z.reshape(-1)


def test_zero_velocity_is_identity():
    s = scene(12, 1.0 / 16)
    rng = np.random.default_rng(1)
    s.phi[:] = rng.uniform(0.1, 1, s.phi.size)
    s.u[:] = s.v[:] = s.w[:] = 0.0
    phi0 = s.phi.copy()
    O.advect_level_set(s, 0.01)
    assert np.array_equal(s.phi, phi0)
    s.u[:] = 0.0
    O.advect_velocity(s, 0.01)
    assert not s.u.any() and not s.v.any() and not s.w.any()


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("cells", [1, 3])
def test_uniform_flow_shifts_by_whole_cells(axis, cells):
    n, dx = 16, 1.0 / 16
    s = scene(n, dx)
    rng = np.random.default_rng(axis)
    s.phi[:] = rng.uniform(0.5, 2.0, s.phi.size)   # no zeros: (1-0) v0 + 0 v1 == v0 exactly
    for a in (s.u, s.v, s.w):
        a[:] = 0.0
    (s.u, s.v, s.w)[axis][:] = 1.0
    phi0 = s.phi.reshape(n, n, n).copy()
    O.advect_level_set(s, cells * dx)
    idx = np.arange(n)
    src = np.clip(idx - cells, 0, n - 1)
    want = np.take(phi0, src, axis=2 - axis)       # array axes are (k, j, i)
    assert np.array_equal(s.phi.reshape(n, n, n), want)


def rotation(s, omega, c):
    n, dx = s.ni, s.dx
    for axis, a in enumerate((s.u, s.v, s.w)):
        x, y, z = face_positions(axis, n, dx)
        if axis == 0:
            a[:] = (-omega * (y - c)).astype(np.float32)
        elif axis == 1:
            a[:] = (omega * (x - c)).astype(np.float32)
        else:
            a[:] = 0.0


def rk2_back(x, y, z, dt, omega, c):
    # A = [[0, -w], [w, 0]] on (x - c, y - c); A^2 = -w^2 I
    rx, ry = x - c, y - c
    ax, ay = -omega * ry, omega * rx
    return (x - dt * ax - 0.5 * dt * dt * omega ** 2 * rx,
            y - dt * ay - 0.5 * dt * dt * omega ** 2 * ry, z)


def test_rotation_level_set_matches_rk2_closed_form():
    n = 32
    dx, dt, omega, c = 1.0 / n, 0.01, 3.0, 0.5
    s = scene(n, dx)
    rotation(s, omega, c)
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    x, y, z = ((i + 0.5) * dx).reshape(-1), ((j + 0.5) * dx).reshape(-1), ((k + 0.5) * dx).reshape(-1)
    a, b, g, d = 0.7, -1.3, 0.4, 0.2
    s.phi[:] = (a * x + b * y + g * z + d).astype(np.float32)
    O.advect_level_set(s, dt)
    bx, by, bz = rk2_back(x, y, z, dt, omega, c)
    want = a * bx + b * by + g * bz + d
    inner = ((i >= 2) & (i < n - 2) & (j >= 2) & (j < n - 2) & (k >= 1) & (k < n - 1)).reshape(-1)
    err = np.abs(s.phi[inner] - want[inner]).max()
    assert err < 2e-6, err
    # RK1 (no dt^2 term) is off by dt^2/2 w^2 |r| |grad| ~ 1e-4 somewhere
    rk1 = a * (x + dt * omega * (y - c)) + b * (y - dt * omega * (x - c)) + g * z + d
    assert np.abs(rk1[inner] - want[inner]).max() > 1e-5


def test_rotation_velocity_self_advection_matches_rk2_closed_form():
    n = 32
    dx, dt, omega, c = 1.0 / n, 0.01, 3.0, 0.5
    s = scene(n, dx)
    rotation(s, omega, c)
    O.advect_velocity(s, dt)
    for axis, a in enumerate((s.u, s.v, s.w)):
        x, y, z = face_positions(axis, n, dx)
        bx, by, bz = rk2_back(x, y, z, dt, omega, c)
        want = [-omega * (by - c), omega * (bx - c), 0.0 * bx][axis]
        dims = [n, n, n]
        dims[axis] += 1
        k, j, i = np.meshgrid(np.arange(dims[2]), np.arange(dims[1]), np.arange(dims[0]), indexing="ij")
        inner = ((i >= 2) & (i < dims[0] - 2) & (j >= 2) & (j < dims[1] - 2)).reshape(-1)
        err = np.abs(a[inner] - want[inner]).max()
        assert err < 2e-6, (axis, err)
