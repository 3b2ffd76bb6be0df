"""The reference's edge semantics on the GPU, checked against the oracle
through the C ABI.

- A solve stopped by max_iter (simulator.cpp:105-109 never checks the
  status): Eigen's loop runs exactly max_iter bodies, iterations() and
  error() report them, and the iterate is the one after the last body.
  FP64 iterates agree with the oracle's to 1e-9 of max|p| (same algebra,
  other summation order of the dots); iterations equal; error() within
  1e-6 relative. max_iter = 0 leaves x = 0, iterations 0, error 1.
- An unknown on the outer cell layer, where build_matrix reads outside its
  arrays (projection.cpp:8-12 via :54): ELQ_ERR_BOUNDARY exactly where the
  oracle's assign_indices reports it, indices bit-exact elsewhere.
- Faces of weight 0 between two unknowns (an interior solid sheet): A holds
  explicit zero couplings (projection.cpp:55-58) and is still assembled bit
  for bit; the solve meets the usual bounds.
- The IC pivots 1/e (the preconditioner set-up, simulator.cpp:102-107):
  bit-exact with the oracle's natural-order IC(0) / MIC(0) factor.
- Status codes: NOT_CONVERGED at max_iter, BREAKDOWN on a non-finite input
  (the Krylov scalars turn NaN; every kernel still ends normally), and a
  context that recovers after a device-side wait ran out (forced from the
  host with a zero wait limit; every wait is bounded, none hangs).
"""
import numpy as np
import pytest

import elq
import oracle as O
from scene_checks import clamped_faces

pytestmark = pytest.mark.gpu


def to_oracle(s):
    return O.Scene(s.ni, s.nj, s.nk, s.dx, s.dt, s.phi.copy(), s.u.copy(), s.v.copy(), s.w.copy(),
                   s.uw.copy(), s.vw.copy(), s.ww.copy())


# ------------------------------------------------------------------ max_iter
@pytest.mark.parametrize("name, max_iter", [("drop32", 0), ("drop32", 5), ("drop32", 50), ("clamp16", 0),
                                            ("clamp16", 5), ("clamp16", 25)])
def test_max_iter_stop_matches_eigen_loop(ctx, name, max_iter):
    # tol 1e-30: the stopping test never fires, every solve runs into max_iter
    s = elq.scene_drop(32) if name == "drop32" else elq.scene_clamp(16)
    g = s.copy()
    o = elq.default_opts(tol=1e-30, max_iter=max_iter, precision=elq.FP64)
    p, idx, st = ctx.solve_pressure(g, o)
    ref = O.solve_pressure(to_oracle(s), tol=1e-30, max_iter=max_iter)
    assert st.status == elq.NOT_CONVERGED and not st.converged, st.as_dict()
    assert st.iterations == ref.iterations == max_iter, (st.iterations, ref.iterations)
    assert abs(st.error - ref.error) <= 1e-6 * ref.error, (st.error, ref.error)
    np.testing.assert_array_equal(idx, ref.indices)
    if max_iter == 0:
        assert st.error == 1.0 and not p.any()
        for f in ("u", "v", "w"):
            np.testing.assert_array_equal(getattr(g, f), getattr(s, f), err_msg=f)
        return
    # FP64 iterates agree to ~1e-12; the float pressure may round one ulp apart
    scale = np.abs(ref.pressure).max()
    ulp = float(np.spacing(np.float32(scale)))
    assert np.abs(p - ref.pressure).max() <= 1e-9 * scale + 2 * ulp, np.abs(p - ref.pressure).max() / scale
    # the update multiplies a pressure difference by at most dt / (dx 0.01)
    du = 4 * ulp * s.dt / (s.dx * 0.01)
    umax = max(np.abs(ref.u).max(), np.abs(ref.v).max(), np.abs(ref.w).max())
    for f in ("u", "v", "w"):
        assert np.abs(getattr(g, f) - getattr(ref, f)).max() <= du + 1e-6 * umax, f


def test_max_iter_bounds_mixed_inner_iterations(ctx):
    s = elq.scene_drop(32)
    _, _, st = ctx.solve_pressure(s.copy(), elq.default_opts(tol=1e-14, max_iter=20, precision=elq.MIXED,
                                                            check_every=4))
    assert st.status == elq.NOT_CONVERGED and not st.converged
    assert st.iterations <= 20 + 4, st.as_dict()


# ------------------------------------------------------------------ outer layer
def _open_box(ni, nj, nk, seed):
    """Face weights 1 everywhere (no room walls), so the outer cell layer is
    valid; the level set is a slab y < 0.45 with random outer-layer cells
    turned to air or liquid and random outer faces closed."""
    rng = np.random.default_rng(seed)
    s = elq.make_scene(ni, nj, nk, 0.45, [])
    for a in (s.uw, s.vw, s.ww):
        a[:] = 1.0
    phi = s.phi.reshape(nk, nj, ni)
    outer = np.zeros((nk, nj, ni), bool)
    outer[[0, -1]] = True
    outer[:, [0, -1]] = True
    outer[:, :, [0, -1]] = True
    if seed % 3 == 0:
        phi[outer] = 1.0                      # no liquid on the outer layer at all
    else:
        flip = outer & (rng.random((nk, nj, ni)) < 0.5)
        phi[flip] = np.where(phi[flip] < 0, np.float32(1.0), np.float32(-1.0))
    if seed % 2 == 1:                         # close the faces of some outer cells: they turn invalid
        u = s.uw.reshape(nk, nj, ni + 1)
        v = s.vw.reshape(nk, nj + 1, ni)
        w = s.ww.reshape(nk + 1, nj, ni)
        shut = outer & (rng.random((nk, nj, ni)) < 0.7)
        for k, j, i in zip(*np.nonzero(shut)):
            u[k, j, i] = u[k, j, i + 1] = 0.0
            v[k, j, i] = v[k, j + 1, i] = 0.0
            w[k, j, i] = w[k + 1, j, i] = 0.0
    rng_v = elq.Rng(seed + 100)
    rng_v.velocity(s, 0.5)
    return s


@pytest.mark.parametrize("seed", range(6))
def test_boundary_error_where_the_oracle_reports_it(ctx, seed):
    s = _open_box(12, 9, 10, seed)
    n_ref, idx_ref = O.assign_indices(to_oracle(s))
    if n_ref < 0:
        with pytest.raises(elq.ElqError) as e:
            ctx.assign_indices(s)
        assert e.value.code == elq.BOUNDARY
        with pytest.raises(elq.ElqError) as e:
            ctx.solve_pressure(s.copy(), elq.default_opts(tol=1e-8))
        assert e.value.code == elq.BOUNDARY
    else:
        n, idx = ctx.assign_indices(s)
        assert n == n_ref
        np.testing.assert_array_equal(idx, idx_ref)


def test_boundary_cases_cover_both_outcomes():
    outcomes = {O.assign_indices(to_oracle(_open_box(12, 9, 10, seed)))[0] < 0 for seed in range(6)}
    assert outcomes == {True, False}


# ------------------------------------------------------------------ explicit zero couplings
def _solid_sheet():
    """Drop scene with two interior solid sheets: u faces at i = 8 for j in
    2..5 and v faces at j = 3 for i in 3..6 get weight 0, between cells that
    stay unknowns (their other faces are open)."""
    s = elq.scene_drop(16)
    n = 16
    u = s.uw.reshape(n, n, n + 1)
    v = s.vw.reshape(n, n + 1, n)
    u[1:15, 2:6, 8] = 0.0
    v[1:15, 3, 3:7] = 0.0
    s.u.reshape(n, n, n + 1)[1:15, 2:6, 8] = 0.0
    s.v.reshape(n, n + 1, n)[1:15, 3, 3:7] = 0.0
    return s


def test_explicit_zero_couplings_bit_exact(ctx):
    s = _solid_sheet()
    os_ = to_oracle(s)
    n_ref, idx_ref = O.assign_indices(os_)
    ref = O.build_matrix(os_, idx_ref, n_ref)
    explicit = int(((ref.xp >= 0) & (ref.ax == 0)).sum() + ((ref.yp >= 0) & (ref.ay == 0)).sum())
    assert explicit >= 40, explicit
    n, idx = ctx.assign_indices(s)
    assert n == n_ref
    np.testing.assert_array_equal(idx, idx_ref)
    m = ctx.build_matrix(s, n)
    for k in ("diag", "ax", "ay", "az", "b"):
        np.testing.assert_array_equal(m[k].view(np.uint64), getattr(ref, k).view(np.uint64), err_msg=k)
    inv_e = ctx.ic_factor(s, n)
    inv_ref, _ = O.ic_factor(ref)
    np.testing.assert_array_equal(inv_e.view(np.uint64), inv_ref.view(np.uint64))


@pytest.mark.parametrize("precision, tol", [(elq.FP64, 2e-6), (elq.MIXED, 2e-5)])
def test_explicit_zero_couplings_solve(ctx, precision, tol):
    s = _solid_sheet()
    g = s.copy()
    p, idx, st = ctx.solve_pressure(g, elq.default_opts(tol=1e-8, max_iter=5000, precision=precision))
    assert st.converged and st.true_rel_residual <= 1e-8, st.as_dict()
    ref = O.solve_pressure(to_oracle(s), tol=1e-12, max_iter=10000)
    np.testing.assert_array_equal(idx, ref.indices)
    assert np.abs(p - ref.pressure).max() <= tol * np.abs(ref.pressure).max()


# ------------------------------------------------------------------ IC pivots
@pytest.mark.parametrize("name", ["drop32", "spheres24", "clamp16", "odd"])
@pytest.mark.parametrize("precond, tau", [(elq.IC0, 0.0), (elq.MIC0, 0.97)])
@pytest.mark.parametrize("schedule", [elq.WAVEFRONT, elq.LEVELS])
def test_ic_pivots_bit_exact(ctx, name, precond, tau, schedule):
    s = {"drop32": lambda: elq.scene_drop(32), "spheres24": lambda: elq.scene_spheres(24, seed=7, n_spheres=6),
         "clamp16": lambda: elq.scene_clamp(16),
         "odd": lambda: _odd()}[name]()
    n, _ = ctx.assign_indices(s)
    inv_e = ctx.ic_factor(s, n, elq.default_opts(precond=precond, tau=tau, schedule=schedule))
    os_ = to_oracle(s)
    n_ref, idx_ref = O.assign_indices(os_)
    ref_inv, fixed = O.ic_factor(O.build_matrix(os_, idx_ref, n_ref), tau, 0.25)
    assert fixed == 0
    np.testing.assert_array_equal(inv_e.view(np.uint64), ref_inv.view(np.uint64))


def _odd():
    s = elq.make_scene(19, 13, 22, 0.45, [(0.5, 0.9, 0.4, 0.2)])
    elq.add_gravity(s, s.dt)
    elq.valid_and_zero(s)
    return s


# ------------------------------------------------------------------ theta clamp scene
def test_clamp_scene_is_clamped_on_every_axis():
    s = elq.scene_clamp(16)
    assert min(clamped_faces(to_oracle(s))) >= 4


# ------------------------------------------------------------------ status codes
@pytest.mark.parametrize("precision", [elq.FP64, elq.MIXED])
def test_breakdown_on_non_finite_input(ctx, precision):
    s = elq.scene_drop(16)
    i, j, k = 8, 3, 8              # a v face inside the slab, between two unknowns
    s.v[i + 16 * (j + 17 * k)] = np.nan  # b, |b|^2 and r.z turn NaN
    p, idx, st = ctx.solve_pressure(s.copy(), elq.default_opts(tol=1e-8, precision=precision), raise_on_error=False)
    assert st.status == elq.BREAKDOWN, st.as_dict()
    # the context is usable afterwards
    t = elq.scene_drop(16)
    ref = O.solve_pressure(to_oracle(t), tol=1e-12, max_iter=10000)
    p, idx, st = ctx.solve_pressure(t, elq.default_opts(tol=1e-8, precision=precision))
    assert st.converged and np.abs(p - ref.pressure).max() <= 2e-5 * np.abs(ref.pressure).max()


@pytest.mark.parametrize("precision", [elq.FP64, elq.MIXED])
def test_context_recovers_after_a_wait_runs_out(precision):
    c = elq.Context(0)
    try:
        s = elq.scene_drop(32)
        c.set_wait_limit(0.0)         # every wait not satisfied at its first check ends the launch
        p, idx, st = c.solve_pressure(s.copy(), elq.default_opts(tol=1e-8, precision=precision),
                                      raise_on_error=False)
        assert st.status == elq.DEVICE, st.as_dict()
        assert "time limit" in c.last_error()
        c.set_wait_limit(4.0)
        ref = O.solve_pressure(to_oracle(s), tol=1e-12, max_iter=10000)
        for _ in range(2):
            p, idx, st = c.solve_pressure(s.copy(), elq.default_opts(tol=1e-8, max_iter=5000, precision=precision))
            assert st.converged and st.true_rel_residual <= 1e-8, st.as_dict()
            np.testing.assert_array_equal(idx, ref.indices)
            assert np.abs(p - ref.pressure).max() <= 2e-5 * np.abs(ref.pressure).max()
        with pytest.raises(elq.ElqError):
            c.set_wait_limit(-1.0)
    finally:
        c.close()
