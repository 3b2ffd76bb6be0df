"""Pins of the CPU oracle (oracle/elq_oracle.c) against facts the mathematics
fixes, since the reference's tests hold no golden vector for the projection
(test/** covers grids and interpolation only) and Eigen, which holds the
reference's solver, is absent (not vendored, not installed).

Each test checks an oracle function against an independent computation:
closed forms, invariants, special cases that reduce to a textbook routine
(dense/sparse direct solves in scipy), or brute force on tiny inputs. A
plausible mistake (a dropped term, a wrong sign or index, a transposed
operand) fails at least one of them; the comment on each test says which.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle as O
from scene_checks import clamped_faces


# ----------------------------------------------------------------- fraction_inside
def test_fraction_inside2_closed_forms():
    # level_set.h:68-80: linear interpolation of the zero crossing
    assert O.fraction_inside2(-1.0, 1.0) == 0.5
    assert O.fraction_inside2(-1.0, -1.0) == 1.0
    assert O.fraction_inside2(1.0, 1.0) == 0.0
    assert O.fraction_inside2(-0.25, 0.75) == np.float32(0.25)
    assert O.fraction_inside2(0.75, -0.25) == np.float32(0.25)  # argument order symmetric
    assert O.fraction_inside2(-0.5, 0.0) == 1.0                 # phi = 0 counts as outside


def _area_linear(bl, br, tl, tr, n=400):
    """Area of {phi < 0} on the unit square for bilinear phi, by sampling."""
    g = (np.arange(n) + 0.5) / n
    x, y = np.meshgrid(g, g)
    phi = bl * (1 - x) * (1 - y) + br * x * (1 - y) + tl * (1 - x) * y + tr * x * y
    return float((phi < 0).mean())


@pytest.mark.parametrize("corners", [(-1, 1, 1, 1), (-1, -1, 1, 1), (-1, 1, -1, 1), (1, -1, -1, -1),
                                     (-0.3, 0.7, -0.2, 0.9), (0.5, -0.5, 0.2, 0.8), (-1, -1, -1, -1),
                                     (1, 1, 1, 1), (-0.01, 1, -0.01, 1)])
def test_fraction_inside4_matches_area_of_linear_level_set(corners):
    # level_set.h:83-163 is exact for a phi whose zero set is a straight line
    # across the cell (affine corner values); corner order bl, br, tl, tr.
    bl, br, tl, tr = corners
    if abs((bl + tr) - (br + tl)) > 1e-6:  # keep only affine corner sets
        tr = br + tl - bl
    got = O.fraction_inside4(bl, br, tl, tr)
    assert abs(got - _area_linear(bl, br, tl, tr)) < 5e-3


def test_fraction_inside4_wall_weight_known_answer():
    # nodal room values -0.01 | +1 along one edge: the negative strip has width
    # 0.01/1.01 of the face, so the open weight is 1 - 0.01/1.01 = 0.990099
    wgt = 1.0 - O.fraction_inside4(-0.01, 1.0, -0.01, 1.0)
    assert abs(wgt - (1 - 0.01 / 1.01)) < 1e-6


# ----------------------------------------------------------------- scene / validity
def test_room_valid_cells_are_interior_box():
    n = 10
    s = O.make_scene(n, n, n, 0.99, [])  # liquid everywhere below y = 0.99
    idx = s.uw.reshape(n, n, n + 1)
    # the two outer u-face layers are closed, deep interior faces fully open
    assert np.all(idx[:, :, :2] == 0) and np.all(idx[:, :, -2:] == 0)
    assert np.all(idx[3:-3, 3:-3, 3:-3] == 1.0)
    nidx, ind = O.assign_indices(s)
    ind = ind.reshape(n, n, n)
    fluid = ind != O.UNKNOWN_INDEX
    expect = np.zeros((n, n, n), bool)
    expect[1:-1, 1:-1, 1:-1] = True
    phi = s.phi.reshape(n, n, n)
    np.testing.assert_array_equal(fluid, expect & (phi < 0))


def test_level_set_closed_forms():
    n = 20
    s = O.make_scene(n, n, n, 0.4, [(0.5, 0.7, 0.5, 0.15)])
    phi = s.phi.reshape(n, n, n)  # [k, j, i]
    c = (np.arange(n) + 0.5) / n
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    expect = np.minimum(np.sqrt((x - 0.5) ** 2 + (y - 0.7) ** 2 + (z - 0.5) ** 2) - 0.15, y - 0.4)
    assert np.abs(phi - expect).max() < 1e-6


def test_assign_indices_is_lexicographic_rank():
    s = O.scene_spheres(14, seed=3, n_spheres=5)
    n, idx = O.assign_indices(s)
    ni = nj = nk = 14
    uw = s.uw.reshape(nk, nj, ni + 1)
    vw = s.vw.reshape(nk, nj + 1, ni)
    ww = s.ww.reshape(nk + 1, nj, ni)
    valid = (uw[:, :, :-1] > 0) | (uw[:, :, 1:] > 0) | (vw[:, :-1, :] > 0) | (vw[:, 1:, :] > 0) | \
            (ww[:-1] > 0) | (ww[1:] > 0)
    fluid = (valid & (s.phi.reshape(nk, nj, ni) < 0)).reshape(-1)
    rank = np.cumsum(fluid) - 1
    expect = np.where(fluid, rank, O.UNKNOWN_INDEX)
    assert n == fluid.sum()
    np.testing.assert_array_equal(idx, expect)


# ----------------------------------------------------------------- assembly
def _system(s):
    n, idx = O.assign_indices(s)
    return n, idx, O.build_matrix(s, idx, n)


def _csr(sys_):
    rp, col, val = O.to_csr(sys_)
    return sp.csr_matrix((val, col, rp), shape=(sys_.n, sys_.n))


def test_matrix_invariants():
    s = O.scene_spheres(16, seed=11, n_spheres=6)
    n, idx, sys_ = _system(s)
    A = _csr(sys_)
    assert abs(A - A.T).max() == 0.0                       # symmetric (same face weight both sides)
    off = A - sp.diags(A.diagonal())
    assert off.max() <= 0.0                                # M-matrix sign pattern
    rowsum = np.asarray(A.sum(axis=1)).ravel()
    assert rowsum.min() >= -1e-12                          # weak diagonal dominance
    # a cell whose six neighbours are fluid through fully open faces: 6 / -1
    k, j, i = 8, 3, 8
    c = i + 16 * (j + 16 * k)
    r = idx[c]
    assert r != O.UNKNOWN_INDEX
    row = A.getrow(r).toarray().ravel()
    if np.count_nonzero(row) == 7:
        assert row[r] == 6.0 and sorted(row[row < 0]) == [-1.0] * 6


@pytest.mark.parametrize("make", [lambda: O.scene_drop(16), lambda: O.scene_spheres(16, seed=11, n_spheres=6)])
def test_ghost_part_of_the_diagonal(make):
    """The FP32 SpMV of the mixed solve stores g = diag + (sum of the row's
    off-diagonals) and computes g p_i + sum a_ij (p_j - p_i). That is the
    same operator as long as g is the row sum of A: g = 0 exactly where all
    six neighbours are unknowns, g > 0 only where a neighbour is not, and the
    off-diagonals are float values (exact in FP32)."""
    s = make()
    n, idx, sys_ = _system(s)
    A = _csr(sys_)
    off = A - sp.diags(A.diagonal())
    nbrs = np.diff(off.tocsr().indptr)
    g = np.asarray(A.sum(axis=1)).ravel()
    assert g.min() >= 0.0
    assert np.all(g[nbrs == 6] == 0.0) and np.count_nonzero(nbrs == 6) > 0
    assert np.all(nbrs[g > 0] < 6) and np.count_nonzero(g > 0) > 0
    vals = off.tocsr().data
    assert np.array_equal(vals.astype(np.float32).astype(np.float64), vals)
    # the difference form gives A p on a smooth field to FP64 rounding
    p = np.linspace(1.0, 2.0, n)
    i, j = off.nonzero()
    q = g * p + np.bincount(i, weights=np.asarray(off[i, j]).ravel() * (p[j] - p[i]), minlength=n)
    assert np.abs(q - A @ p).max() <= 1e-12 * np.abs(A).max() * np.abs(p).max()


def _divergence(s):
    """Weighted divergence sum_faces (+/-) w u / dx, cell-centred, [k, j, i]."""
    ni, nj, nk = s.ni, s.nj, s.nk
    u = (s.u * s.uw).astype(np.float64).reshape(nk, nj, ni + 1)
    v = (s.v * s.vw).astype(np.float64).reshape(nk, nj + 1, ni)
    w = (s.w * s.ww).astype(np.float64).reshape(nk + 1, nj, ni)
    return ((u[:, :, 1:] - u[:, :, :-1]) + (v[:, 1:, :] - v[:, :-1, :]) + (w[1:] - w[:-1])) / s.dx


def test_rhs_is_scaled_negative_divergence():
    # projection.cpp:65-69,82: b = (dx/dt) * sum(-w u_fwd + w u_bwd) = -(dx^2/dt) div
    s = O.scene_spheres(16, seed=5, n_spheres=6)
    n, idx, sys_ = _system(s)
    div = _divergence(s).reshape(-1)
    fluid = idx != O.UNKNOWN_INDEX
    expect = -(np.float32(s.dx) / np.float32(s.dt)) * s.dx * div[fluid]
    got = sys_.b[idx[fluid]]
    assert np.abs(got - expect).max() <= 1e-6 * np.abs(expect).max()


def test_uniform_flow_has_zero_rhs_in_open_cells():
    s = O.make_scene(12, 12, 12, 0.8, [])
    s.u[:] = 0.3
    s.v[:] = -0.2
    s.w[:] = 0.1
    n, idx, sys_ = _system(s)
    idx3 = idx.reshape(12, 12, 12)
    for (k, j, i) in [(5, 5, 5), (6, 4, 7)]:
        assert sys_.b[idx3[k, j, i]] == 0.0


# ----------------------------------------------------------------- ghost-fluid theta clamp
# projection.cpp:59-63 (assembly) and :100-131 (update): theta =
# max(fraction_inside(phi_c, phi_n), 0.01f). An interface that passes close to
# a cell centre gives theta ~ 1e-4; the clamp turns it into 0.01. The known
# answers below are worked out by hand from those lines (w = 1, float 1/0.01f
# rounds to exactly 100): an oracle without the clamp (w/theta ~ 1e4 per
# face), with min in place of max (the theta = 0.5 face would give 100, not
# 2) or with another constant fails them.
def _isolated_cell(axis, up):
    """8^3 room scene, all air (phi = 1) except cell (3,3,3) at phi = -1e-4;
    its neighbour across face (axis, up) sits at phi = +1e-4 (theta = 0.5
    exactly), the other five at phi = 1 (theta ~ 1e-4, clamped)."""
    n = 8
    s = O.make_scene(n, n, n, None, [])
    phi = np.ones((n, n, n), np.float32)          # [k, j, i]
    phi[3, 3, 3] = np.float32(-1e-4)
    nb = [3, 3, 3]
    nb[2 - axis] += 1 if up else -1
    phi[tuple(nb)] = np.float32(1e-4)
    s.phi[:] = phi.reshape(-1)
    return s


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("up", [False, True])
def test_theta_clamp_known_answer_in_assembly(axis, up):
    s = _isolated_cell(axis, up)
    assert s.uw.reshape(8, 8, 9)[3, 3, 3] == 1.0   # room(): interior faces fully open
    n, idx, sys_ = _system(s)
    assert n == 1
    # five clamped faces at w/0.01 = 100 and one at w/0.5 = 2
    assert sys_.diag[0] == 502.0, sys_.diag[0]


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_theta_clamp_known_answer_in_update(axis):
    # pressure 1 in the fluid cell, 0 in the air; u -= dt (p_hi - p_lo) / (dx theta):
    # the -axis face (towards phi = 1, theta clamped to 0.01) moves by
    # -0.01 * 1 / (0.125 * 0.01) = -8, the +axis face (towards phi = +1e-4,
    # theta = 0.5) by -0.01 * (0 - 1) / (0.125 * 0.5) = +0.16
    s = _isolated_cell(axis, True)
    p = np.zeros(512, np.float32)
    p[3 + 8 * (3 + 8 * 3)] = 1.0
    grids = (s.u, s.v, s.w)
    shapes = ((8, 8, 9), (8, 9, 8), (9, 8, 8))
    before = [g.copy() for g in grids]
    O.update_velocity(s, p)
    g = (grids[axis] - before[axis]).reshape(shapes[axis])
    lo = [3, 3, 3]
    hi = [3, 3, 3]
    hi[2 - axis] += 1                              # face index of the +axis face
    assert abs(g[tuple(lo)] - (-8.0)) <= 1e-5 * 8.0, g[tuple(lo)]   # -axis face: theta clamped
    assert abs(g[tuple(hi)] - 0.16) <= 1e-5 * 0.16, g[tuple(hi)]    # +axis face: theta = 0.5


def test_clamp_scene_has_clamped_faces_on_every_axis():
    s = O.scene_clamp(16)
    assert min(clamped_faces(s)) >= 4, clamped_faces(s)


# ----------------------------------------------------------------- preconditioner
def _ic_matrix(sys_, inv_e):
    A = _csr(sys_)
    L = sp.tril(A, k=-1).tocsr()
    E = sp.diags(1.0 / inv_e)
    return (E + L) @ sp.diags(inv_e) @ (E + L.T), A


def test_ic0_reproduces_A_on_its_pattern():
    # IC(0) defining property: (E+L)E^-1(E+L^T) equals A wherever A has an entry
    s = O.scene_spheres(12, seed=2, n_spheres=4)
    n, idx, sys_ = _system(s)
    inv_e, fixed = O.ic_factor(sys_, 0.0, 0.25)
    assert fixed == 0
    M, A = _ic_matrix(sys_, inv_e)
    pattern = A.copy()
    pattern.data[:] = 1.0
    diff = (M - A).multiply(pattern)
    assert abs(diff).max() <= 1e-12 * abs(A).max()
    assert abs(M - A).max() > 1e-6  # but the dropped fill is non-zero


def test_mic0_preserves_row_sums():
    # MIC(0) with tau = 1: dropped fill is added to the diagonal, M 1 = A 1
    s = O.scene_spheres(12, seed=2, n_spheres=4)
    n, idx, sys_ = _system(s)
    inv_e, fixed = O.ic_factor(sys_, 1.0, 0.0)
    M, A = _ic_matrix(sys_, inv_e)
    one = np.ones(n)
    assert np.abs(M @ one - A @ one).max() <= 1e-10 * abs(A).max()


def test_ic_apply_inverts_M():
    s = O.scene_spheres(12, seed=9, n_spheres=4)
    n, idx, sys_ = _system(s)
    inv_e, _ = O.ic_factor(sys_)
    M, _ = _ic_matrix(sys_, inv_e)
    r = np.random.default_rng(0).standard_normal(n)
    z = O.ic_apply(sys_, inv_e, r)
    assert np.abs(M @ z - r).max() <= 1e-10 * np.abs(r).max()


def test_single_column_is_exact_cholesky():
    # a 1 x n x 1 column of fluid is tridiagonal: IC(0) drops no fill, M = A,
    # and preconditioned CG converges in one iteration
    n = 10
    s = O.make_scene(n, n, n, 0.7, [])
    keep = np.zeros((n, n, n), bool)
    keep[5, :, 5] = True
    s.phi[~keep.reshape(-1)] = 1.0
    s.v[:] = np.random.default_rng(1).standard_normal(s.v.size).astype(np.float32)
    nunk, idx, sys_ = _system(s)
    assert nunk == int((s.phi < 0).sum()) - 1  # the j = 0 cell lies on the closed outer layer
    inv_e, _ = O.ic_factor(sys_)
    M, A = _ic_matrix(sys_, inv_e)
    assert abs(M - A).max() <= 1e-12 * abs(A).max()
    x, it, err = O.pcg(sys_, inv_e, tol=1e-12, max_iter=10)
    assert it <= 1


# ----------------------------------------------------------------- PCG (Eigen semantics)
@pytest.mark.parametrize("seed", [1, 4])
def test_pcg_matches_direct_solve(seed):
    s = O.scene_spheres(14, seed=seed, n_spheres=5)
    n, idx, sys_ = _system(s)
    inv_e, _ = O.ic_factor(sys_)
    x, it, err = O.pcg(sys_, inv_e, tol=1e-13, max_iter=2000)
    A = _csr(sys_).tocsc()
    xd = spla.spsolve(A, sys_.b)
    assert np.abs(x - xd).max() <= 1e-9 * np.abs(xd).max()
    assert err < 1e-13
    r = sys_.b - A @ x
    assert np.linalg.norm(r) / np.linalg.norm(sys_.b) < 1e-12


def test_pcg_manufactured_solution():
    s = O.scene_spheres(12, seed=8, n_spheres=3)
    n, idx, sys_ = _system(s)
    pstar = np.random.default_rng(3).standard_normal(n)
    sys_.b[:] = O.spmv(sys_, pstar)
    assert np.allclose(sys_.b, _csr(sys_) @ pstar, rtol=0, atol=1e-12)
    inv_e, _ = O.ic_factor(sys_)
    x, it, err = O.pcg(sys_, inv_e, tol=1e-14, max_iter=5000)
    assert np.abs(x - pstar).max() < 1e-8


def test_pcg_eigen_edge_semantics():
    s = O.scene_spheres(10, seed=2, n_spheres=2)
    n, idx, sys_ = _system(s)
    inv_e, _ = O.ic_factor(sys_)
    b = sys_.b.copy()
    sys_.b[:] = 0.0
    x, it, err = O.pcg(sys_, inv_e, tol=1e-4)
    assert it == 0 and err == 0.0 and not x.any()        # |b| = 0: x = 0
    sys_.b[:] = b
    x, it, err = O.pcg(sys_, inv_e, tol=2.0)
    assert it == 0 and abs(err - 1.0) < 1e-15             # |r0| = |b| < tol |b|: no iteration
    x, it, err = O.pcg(sys_, inv_e, tol=1e-12, max_iter=1)
    assert it == 1 and err > 1e-12                        # max_iter reached, error reported
    x, it, err = O.pcg(sys_, inv_e, tol=1e-4, max_iter=1000)
    assert err < 1e-4 and it >= 1


# ----------------------------------------------------------------- update (projection.cpp:93-134)
def test_projection_divergence_identity():
    # after the update, in every fluid cell:
    #   div' = -(dt/dx^2) (b - A p_float)  up to float rounding of the update
    s = O.scene_spheres(16, seed=6, n_spheres=6)
    n, idx, sys_ = _system(s)
    res = O.solve_pressure(s, tol=1e-12, max_iter=5000)
    t = s.copy()
    t.u, t.v, t.w = res.u, res.v, res.w
    div_after = _divergence(t).reshape(-1)
    fluid = idx != O.UNKNOWN_INDEX
    p = res.pressure.astype(np.float64)[fluid][np.argsort(idx[fluid])]
    A = _csr(sys_)
    expect = -(s.dt / s.dx ** 2) * (sys_.b - A @ p)
    got = div_after[fluid][np.argsort(idx[fluid])]
    div_before = np.abs(_divergence(s).reshape(-1)[fluid]).max()
    assert np.abs(got - expect).max() <= 1e-5 * div_before
    assert np.abs(got).max() <= 1e-4 * div_before        # the projection removed the divergence


def test_update_skips_boundary_tangential_rows():
    # projection.cpp:98-99,110-111,122-123: ranges {1, n_face - 1}
    s = O.make_scene(8, 8, 8, 0.9, [])
    s.phi[:] = -1.0
    u0 = s.u.copy()
    p = np.random.default_rng(0).standard_normal(s.ni * s.nj * s.nk).astype(np.float32)
    O.update_velocity(s, p)
    u = s.u.reshape(8, 8, 9)
    du = (u - u0.reshape(8, 8, 9)) != 0
    assert not du[:, :, 0].any() and not du[:, :, 8].any()   # normal index 0 and ni untouched
    assert not du[0].any() and not du[7].any()               # tangential k rows 0 and nk-1
    assert not du[:, 0].any() and not du[:, 7].any()         # tangential j rows 0 and nj-1
    assert du[1:7, 1:7, 1:8].all()


# ----------------------------------------------------------------- scene inputs
def test_splitmix64_known_answers():
    # published test vector of SplitMix64 (Steele, Lea, Flood 2014; Vigna's
    # reference implementation), seed 1234567: a wrong constant or shift fails
    import ctypes as C
    st = C.c_uint64(1234567)
    got = [O.lib().orc_splitmix64(C.byref(st)) for _ in range(5)]
    assert got == [6457827717110365317, 3203168211198807973, 9817491932198370423, 4593380528125082431,
                   16408922859458223821]
    st = C.c_uint64(1234567)
    u = O.lib().orc_uniform01(C.byref(st))
    assert u == (6457827717110365317 >> 11) * 2.0 ** -53    # top 53 bits, [0, 1)


def test_add_gravity_uses_reference_constant_on_every_v_face():
    # simulator.cpp:90-93 with constants.h:5-6 (9.80665): all v faces, u and w untouched
    s = O.make_scene(6, 5, 4, 0.5, [])
    O.add_gravity(s, 0.01)
    assert np.all(s.v == np.float32(0.0) - np.float32(9.80665) * np.float32(0.01))
    assert not s.u.any() and not s.w.any()


def _ones_scene(ni, nj, nk, phi_value):
    s = O.make_scene(ni, nj, nk, 0.5, [])
    s.phi[:] = phi_value
    for a in (s.uw, s.vw, s.ww, s.u, s.v, s.w):
        a[:] = 1.0
    return s


def test_valid_velocity_special_cases():
    # simulator.cpp:125-178. All liquid, all weights 1: nothing is zeroed.
    s = _ones_scene(6, 5, 4, -1.0)
    O.find_valid_and_zero(s)
    assert s.u.all() and s.v.all() and s.w.all()
    # weight 0 alone zeroes a face
    s = _ones_scene(6, 5, 4, -1.0)
    s.uw[7] = 0.0
    O.find_valid_and_zero(s)
    assert s.u[7] == 0.0 and np.count_nonzero(s.u == 0.0) == 1
    # no liquid: exactly the faces inside the loop ranges are zeroed; ranges
    # are [1, n_face - 1) per axis, so the last index of each axis survives
    ni, nj, nk = 6, 5, 4
    s = _ones_scene(ni, nj, nk, 1.0)
    O.find_valid_and_zero(s)
    u = s.u.reshape(nk, nj, ni + 1)
    keep = np.ones_like(u, bool)
    keep[1:nk - 1, 1:nj - 1, 1:ni] = False
    assert np.array_equal(u != 0.0, keep)
    v = s.v.reshape(nk, nj + 1, ni)
    keep = np.ones_like(v, bool)
    keep[1:nk - 1, 1:nj, 1:ni - 1] = False
    assert np.array_equal(v != 0.0, keep)
    w = s.w.reshape(nk + 1, nj, ni)
    keep = np.ones_like(w, bool)
    keep[1:nk, 1:nj - 1, 1:ni - 1] = False
    assert np.array_equal(w != 0.0, keep)
    # one liquid cell: its six faces survive besides the untouched rows
    s = _ones_scene(ni, nj, nk, 1.0)
    c = (2, 2, 1)  # i, j, k
    s.phi[c[0] + ni * (c[1] + nj * c[2])] = -1.0
    O.find_valid_and_zero(s)
    u = s.u.reshape(nk, nj, ni + 1)
    assert u[1, 2, 2] == 1.0 and u[1, 2, 3] == 1.0 and u[1, 2, 4] == 0.0
    v = s.v.reshape(nk, nj + 1, ni)
    assert v[1, 2, 2] == 1.0 and v[1, 3, 2] == 1.0 and v[1, 1, 2] == 0.0
    w = s.w.reshape(nk + 1, nj, ni)
    assert w[1, 2, 2] == 1.0 and w[2, 2, 2] == 1.0 and w[3, 2, 2] == 0.0


def test_build_grid_from_vector_scatter():
    # projection.cpp:85-91: pressure(c) = (float)x[idx(c)], background 0
    s = O.scene_drop(12)
    n, idx = O.assign_indices(s)
    x = np.arange(n, dtype=np.float64) * (1.0 + 2.0 ** -40) + 0.25
    p = O.build_grid_from_vector(s, x, idx)
    fluid = idx != O.UNKNOWN_INDEX
    assert np.array_equal(p[fluid], x[idx[fluid]].astype(np.float32))
    assert not p[~fluid].any()
