"""Pins of the oracle's extrapolation and validity flags (extrapolation.cpp:5-51,
simulator.cpp:125-200) against facts fixed without the oracle's own formula:

- the set of cells valid after L layers is the valid set dilated L times by
  the 3x3x3 cube, i.e. the cells within Chebyshev distance L (a 6-neighbour
  stencil would give an octahedron, one layer too many or too few a ball of
  the wrong radius, a layer that reads values written in the same layer a
  ball of radius > L);
- means of small integers, computed by hand, for cells that see 1, 2 and 4
  valid neighbours;
- the float summation order: 1e8, -1e8 and 1 summed in the reference's order
  (k slowest, then j, then i) give 1; with i slowest (a transposed loop
  nest) they give 0;
- cells never reached keep their old values (not zeroed); all-valid and
  no-valid grids are unchanged; extrapolate_const leaves the flags alone;
- find_valid_level_set on room(): the cells whose 8 nodes average to < 0
  are exactly the outer layer.
"""
import numpy as np
import pytest

import oracle as O


def cheb_ball(dims, src, r):
    ni, nj, nk = dims
    z, y, x = np.meshgrid(np.arange(nk), np.arange(nj), np.arange(ni), indexing="ij")
    d = np.full((nk, nj, ni), np.iinfo(np.int32).max)
    for (i, j, k) in src:
        d = np.minimum(d, np.maximum(np.maximum(abs(x - i), abs(y - j)), abs(z - k)))
    return d <= r


@pytest.mark.parametrize("src,layers", [((7, 6, 5), 4), ((0, 0, 0), 3), ((14, 1, 10), 6), ((7, 6, 5), 1)])
def test_single_source_fills_chebyshev_ball(src, layers):
    dims = (15, 13, 11)
    ni, nj, nk = dims
    rng = np.random.default_rng(1)
    g = rng.uniform(-5, 5, ni * nj * nk).astype(np.float32)
    old = g.copy()
    v = np.zeros(g.size, np.uint8)
    c = src[0] + ni * (src[1] + nj * src[2])
    g[c] = 3.0
    v[c] = 1
    O.extrapolate(g, v, dims, layers)
    ball = cheb_ball(dims, [src], layers).reshape(-1)
    assert np.array_equal(v.astype(bool), ball)
    assert np.all(g[ball] == np.float32(3.0))          # mean of equal small integers is exact
    assert np.array_equal(g[~ball], old[~ball])         # unreached cells keep their values


def test_hand_computed_means_and_layer_snapshot():
    dims = (10, 10, 10)
    n = 10
    idx = lambda i, j, k: i + n * (j + n * k)
    g = np.zeros(n ** 3, np.float32)
    v = np.zeros(n ** 3, np.uint8)
    val = {}
    q = 1
    for k in (4, 5):
        for j in (4, 5):
            for i in (4, 5):
                g[idx(i, j, k)] = q
                v[idx(i, j, k)] = 1
                val[(i, j, k)] = q
                q += 1
    g1, v1 = g.copy(), v.copy()
    O.extrapolate(g1, v1, dims, 1)
    # corner-diagonal cell: one valid neighbour
    assert g1[idx(3, 3, 3)] == val[(4, 4, 4)]
    # edge cell (3,3,4): valid neighbours (4,4,4) and (4,4,5)
    assert g1[idx(3, 3, 4)] == (val[(4, 4, 4)] + val[(4, 4, 5)]) / 2
    # face cell (3,4,4): the four cells with i = 4
    face = [val[(4, j, k)] for k in (4, 5) for j in (4, 5)]
    assert g1[idx(3, 4, 4)] == np.float32(sum(face) / 4)
    # a cell at distance 2 is untouched after one layer
    assert v1[idx(2, 2, 2)] == 0 and g1[idx(2, 2, 2)] == 0
    # after two layers (2,2,2) sees only (3,3,3), whose value came from (4,4,4)
    g2, v2 = g.copy(), v.copy()
    O.extrapolate(g2, v2, dims, 2)
    assert v2[idx(2, 2, 2)] == 1 and g2[idx(2, 2, 2)] == val[(4, 4, 4)]
    # (2,3,3) in layer 2 sees the four layer-1 cells (3, j, k), j, k in {3, 4};
    # they are summed k slowest, then j
    nb = [g1[idx(3, j, k)] for k in (2, 3, 4) for j in (2, 3, 4) if v1[idx(3, j, k)]]
    assert len(nb) == 4
    want = np.float32(0)
    for x in nb:
        want = np.float32(want + x)
    assert g2[idx(2, 3, 3)] == np.float32(want / np.float32(len(nb)))


def test_summation_order_is_k_then_j_then_i():
    dims = (5, 5, 5)
    n = 5
    idx = lambda i, j, k: i + n * (j + n * k)
    g = np.zeros(n ** 3, np.float32)
    v = np.zeros(n ** 3, np.uint8)
    c = (2, 2, 2)
    # three valid neighbours of c, listed in the reference's order (k, then j, then i)
    cells = [((3, 2, 1), 1e8), ((2, 1, 2), -1e8), ((1, 2, 3), 1.0)]
    for (i, j, k), x in cells:
        g[idx(i, j, k)] = x
        v[idx(i, j, k)] = 1
    O.extrapolate(g, v, dims, 1)
    # (1e8 + -1e8) + 1 = 1 -> 1/3; with i slowest (a transposed loop nest) the
    # order is 1, -1e8, 1e8 and the float sum is 0
    assert g[idx(*c)] == np.float32(1.0) / np.float32(3.0)


def test_trivial_grids_and_const_variant():
    dims = (7, 6, 5)
    rng = np.random.default_rng(2)
    g = rng.uniform(-1, 1, 210).astype(np.float32)
    v = np.ones(210, np.uint8)
    g0 = g.copy()
    O.extrapolate(g, v, dims, 10)          # all valid: early return, nothing changes
    assert np.array_equal(g, g0) and v.all()
    v = np.zeros(210, np.uint8)
    O.extrapolate(g, v, dims, 10)          # nothing valid: nothing changes
    assert np.array_equal(g, g0) and not v.any()
    v = (rng.random(210) < 0.05).astype(np.uint8)
    v0 = v.copy()
    a, b = g.copy(), g.copy()
    va = v.copy()
    O.extrapolate(a, va, dims, 2)
    O.extrapolate_const(b, v, dims, 2)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert np.array_equal(v, v0) and not np.array_equal(va, v0)


def test_level_set_valid_on_room():
    for (ni, nj, nk) in ((8, 7, 6), (12, 12, 12)):
        nodal = np.empty((ni + 1) * (nj + 1) * (nk + 1), np.float32)
        O.lib().orc_room_nodal(ni, nj, nk, nodal)
        v = np.empty(ni * nj * nk, np.uint8)
        O.lib().orc_find_valid_level_set(ni, nj, nk, nodal, v)
        # room() on the nodal grid: nodes on layer 0 / n are -1, on 1 / n-1
        # -0.01, else 1. An outer-layer cell has at least 4 nodes at -1 and the
        # others at -0.01: mean < 0. Every other cell has at least one node at
        # 1 and the rest at -0.01 or 1: mean > 0.
        inner = np.zeros((nk, nj, ni), bool)
        inner[1:nk - 1, 1:nj - 1, 1:ni - 1] = True
        assert np.array_equal(v.reshape(nk, nj, ni).astype(bool), inner)


def test_find_valid_velocity_flags_match_zeroing():
    # the flags and the zeroing agree face by face on a scene with liquid and air
    s = O.make_scene(12, 10, 9, 0.4, [(0.5, 0.7, 0.5, 0.2)])
    for a in (s.u, s.v, s.w):
        a[:] = 1.0
    flags = O.find_valid_velocity(s)
    O.find_valid_and_zero(s)
    for a, f in zip((s.u, s.v, s.w), flags):
        assert np.array_equal(a != 0, f.astype(bool))
        assert f.any() and not f.all()
