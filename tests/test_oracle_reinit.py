"""Pins of the oracle's reinitialisation (reinitialization.cpp:19-156 with
level_set.h:165-264 and meshing.h:12-17) against facts fixed without the
oracle's formulas:

- first-order upwind marching is exact for planar fronts: a level set
  k (n . x - c) with any scale k becomes the signed distance (n . x - c)
  for n along an axis (one upwind neighbour: d = p + dx), at 45 degrees in a
  coordinate plane (two neighbours: the quadratic with sqrt(2 dx^2 - ...))
  and along (1,1,1)/sqrt(3) (three neighbours; the reference's 0.333333 for
  1/3 costs about 1e-6 relative per pass). The interface cells come from the
  geometric distance to the interpolated zero set, exact for a plane. Cells
  checked are those whose upwind cone stays off the domain walls. A wrong
  formula, a neighbour taken from the wrong side, a wrong tetrahedron or a
  missing sign flip misses by a sizeable fraction of dx;
- the three-neighbour update is 0.999999 times the exact root of its
  upwind quadratic (the reference's 0.333333 for 1/3);
- signs are kept; invalid cells are untouched; cells already at 0 stay 0;
- a pocket of cells walled off by invalid cells is reported (the reference
  would loop forever there) and left unchanged.
"""
import numpy as np
import pytest

import oracle as O


def grid_xyz(n, dx):
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    return (i + 0.5) * dx, (j + 0.5) * dx, (k + 0.5) * dx, i, j, k


@pytest.mark.parametrize("normal,scale,tol", [((0, 1, 0), 3.0, 2e-6), ((1, 0, 0), 0.2, 2e-6),
                                              ((1, 1, 0), 2.0, 2e-6), ((0, 1, -1), 0.5, 2e-6),
                                              ((1, 1, 1), 1.7, 2e-5), ((-1, 1, 1), 4.0, 2e-5)])
def test_planar_fronts_become_signed_distances(normal, scale, tol):
    n, dx = 40, 1.0 / 40
    x, y, z, i, j, k = grid_xyz(n, dx)
    nv = np.array(normal, float) / np.linalg.norm(normal)
    c = nv @ np.array([0.5, 0.5, 0.5]) + 0.3 * dx          # plane off the cell centres
    dist = nv[0] * x + nv[1] * y + nv[2] * z - c
    phi = (scale * dist).astype(np.float32).reshape(-1)
    valid = np.ones(phi.size, np.uint8)
    left, passes = O.reinitialize(phi, valid, (n, n, n), dx)
    assert left == 0 and passes > 10
    layers = np.abs(dist) / dx * np.abs(nv).sum()            # upwind cone depth in cells
    m = np.ceil(layers) + 2
    inner = (np.minimum(np.minimum(i, j), k) >= m) & (np.maximum(np.maximum(i, j), k) <= n - 1 - m)
    inner = inner.reshape(-1)
    assert inner.sum() > 500
    err = np.abs(phi[inner] - dist.reshape(-1)[inner]).max()
    assert err < tol, err
    assert np.array_equal(np.sign(phi), np.sign(dist.reshape(-1)))


def test_three_neighbour_update_uses_the_references_third():
    # reinitialization.cpp writes 1/3 as the literal 0.333333, so its
    # three-neighbour update is 0.999999 times the exact root of the upwind
    # quadratic (sum_a (d - p_a)^2 = dx^2) of the FIX values it reads. On the
    # positive side of a (1,1,1) plane, the cells of the first marched layer
    # whose -x, -y and -z neighbours are interface cells take that update;
    # the ratio of their value to the exact root of their neighbours' values
    # is 0.999999 up to float rounding (an exact third would give 1).
    n, dx = 40, 1.0 / 32
    x, y, z, i, j, k = grid_xyz(n, dx)
    nv = np.ones(3) / np.sqrt(3.0)
    c = nv @ np.array([0.6, 0.6, 0.6]) + 0.3 * dx
    phi = (nv[0] * x + nv[1] * y + nv[2] * z - c).astype(np.float32).reshape(-1)
    sgn = np.sign(phi).reshape(n, n, n)
    iface = np.zeros((n, n, n), bool)
    for ax in range(3):
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[ax], hi[ax] = slice(0, -1), slice(1, None)
        flip = sgn[tuple(lo)] != sgn[tuple(hi)]
        iface[tuple(lo)] |= flip
        iface[tuple(hi)] |= flip
    valid = np.ones(phi.size, np.uint8)
    O.reinitialize(phi, valid, (n, n, n), dx)
    v = phi.reshape(n, n, n).astype(np.float64)
    sel = (sgn > 0) & ~iface
    sel[0, :, :] = sel[:, 0, :] = sel[:, :, 0] = False
    sel[1:, 1:, 1:] &= iface[:-1, 1:, 1:] & iface[1:, :-1, 1:] & iface[1:, 1:, :-1]
    kk, jj, ii = np.nonzero(sel)
    assert kk.size > 200
    p = np.stack([v[kk, jj, ii - 1], v[kk, jj - 1, ii], v[kk - 1, jj, ii]])
    s_ = p.sum(0)
    root = (s_ + np.sqrt(s_ * s_ - 3.0 * ((p * p).sum(0) - dx * dx))) / 3.0
    assert np.all(root > p.max(0))        # the three-neighbour case applies
    ratio = v[kk, jj, ii] / root
    assert abs(np.median(ratio) - 0.999999) < 2e-7, np.median(ratio)
    assert np.abs(ratio - 0.999999).max() < 8e-7


def test_invalid_cells_zero_cells_and_unreachable_pockets():
    n, dx = 24, 1.0 / 24
    x, y, z, i, j, k = grid_xyz(n, dx)
    phi = (y - 0.5 - 0.2 * dx).astype(np.float32).reshape(-1)
    phi[(i == 5).reshape(-1) & (j == 2).reshape(-1) & (k == 7).reshape(-1)] = 0.0
    valid = np.ones(phi.size, np.uint8)
    # a 3^3 pocket around (18, 4, 18) (all phi < 0) walled off by a shell of invalid cells
    shell = (np.abs(i - 18) <= 2) & (np.abs(j - 4) <= 2) & (np.abs(k - 18) <= 2)
    pocket = (np.abs(i - 18) <= 1) & (np.abs(j - 4) <= 1) & (np.abs(k - 18) <= 1)
    valid[(shell & ~pocket).reshape(-1)] = 0
    valid[(i == 0).reshape(-1)] = 0
    phi0 = phi.copy()
    left, passes = O.reinitialize(phi, valid, (n, n, n), dx)
    assert left == 27
    inval = valid == 0
    assert np.array_equal(phi[inval], phi0[inval])
    assert np.array_equal(phi[pocket.reshape(-1)], phi0[pocket.reshape(-1)])
    zero = phi0 == 0.0
    assert zero.sum() == 1 and np.all(phi[zero] == 0.0)
    ok = (valid != 0) & ~pocket.reshape(-1) & ~zero
    assert np.array_equal(np.sign(phi[ok]), np.sign(phi0[ok]))


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_meeting_fronts_take_the_lower_neighbour(axis):
    # two parallel interfaces; their fronts meet at a ridge cell whose two
    # neighbours along the axis are both fixed in the same pass with
    # different values: the reference takes the one on the lower side
    # (reinitialization.cpp:83-97 checks -x before +x, -y before +y, ...)
    n, dx = 32, 1.0 / 32
    x, y, z, i, j, k = grid_xyz(n, dx)
    t = (x, y, z)[axis]
    a, b = 0.3 + 0.1 * dx, 0.7 + 0.35 * dx
    phi = (2.0 * np.minimum(t - a, b - t)).astype(np.float32).reshape(-1)
    valid = np.ones(phi.size, np.uint8)
    left, _ = O.reinitialize(phi, valid, (n, n, n), dx)
    assert left == 0
    A = phi.reshape(n, n, n)                                 # [k, j, i]
    line = (A[16, 16, :], A[16, :, 16], A[:, 16, 16])[axis]
    inside = np.where(line > 0)[0]
    ridge = [q for q in inside[1:-1] if line[q - 1] < line[q] and line[q + 1] < line[q]]
    assert len(ridge) == 1
    q = ridge[0]
    assert line[q - 1] != line[q + 1]
    assert line[q] == np.float32(line[q - 1] + np.float32(dx))
