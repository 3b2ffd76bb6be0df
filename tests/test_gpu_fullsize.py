"""Projection at the benchmark sizes, through the C ABI, checked by properties
that do not need a CPU solve of the same size, and at 128^3 against the oracle.

- 256^3 slab + 32 spheres (the benchmark scene): mixed precision and the FP64
  drop-in both reach a true FP64 relative residual <= 1e-8; their pressures
  agree to max|p_mixed - p_FP64| <= 1e-5 max|p|; after the velocity update the
  weighted divergence of every fluid cell is <= 1e-4 of the largest divergence
  before the projection (the identity div' = -(dt/dx^2)(b - A p) turns the
  1e-8 residual into a divergence bound; 1e-4 leaves room for the float
  pressure and update).
- 128^3 drop scene: pressure within 2e-6 max|p| of the oracle's FP64 solve to
  relative residual 1e-12 (mixed: 2e-5).
- 512^3 slab + 128 spheres (about 48.7 M unknowns, past what a test can solve
  on the CPU): the same properties as at 256^3 on device-resident grids, and
  both precisions converge within the reference's 1000 iterations
  (simulator.cpp:105); cell and unknown indices stay below 2^31 (the
  reference's int32 guard, array.h:322-325).
"""
import numpy as np
import pytest

import elq
import oracle as O

pytestmark = pytest.mark.gpu


def divergence(s, u, v, w):
    ni, nj, nk = s.ni, s.nj, s.nk
    uu = (u * s.uw).astype(np.float64).reshape(nk, nj, ni + 1)
    vv = (v * s.vw).astype(np.float64).reshape(nk, nj + 1, ni)
    ww = (w * s.ww).astype(np.float64).reshape(nk + 1, nj, ni)
    return ((uu[:, :, 1:] - uu[:, :, :-1]) + (vv[:, 1:, :] - vv[:, :-1, :]) + (ww[1:] - ww[:-1])) / s.dx


@pytest.fixture(scope="module")
def scene256():
    return elq.scene_spheres(256, seed=1, n_spheres=32)


def test_256_mixed_and_fp64(ctx, scene256):
    s = scene256
    div0 = np.abs(divergence(s, s.u, s.v, s.w)).max()
    out = {}
    for prec in (elq.MIXED, elq.FP64):
        t = s.copy()
        p, idx, st = ctx.solve_pressure(t, elq.default_opts(tol=1e-8, max_iter=20000, precision=prec))
        assert st.converged and st.status == elq.OK, st.as_dict()
        assert st.true_rel_residual <= 1e-8, st.true_rel_residual
        assert 100 < st.iterations < 1000
        fluid = (idx != elq.UNKNOWN_INDEX).reshape(s.nk, s.nj, s.ni)
        div = divergence(t, t.u, t.v, t.w)
        assert np.abs(div[fluid]).max() <= 1e-4 * div0
        out[prec] = p
    scale = np.abs(out[elq.FP64]).max()
    assert np.abs(out[elq.MIXED] - out[elq.FP64]).max() <= 1e-5 * scale


@pytest.fixture(scope="module")
def drop128():
    s = elq.scene_drop(128)
    ref = O.solve_pressure(O.Scene(s.ni, s.nj, s.nk, s.dx, s.dt, s.phi.copy(), s.u.copy(), s.v.copy(), s.w.copy(),
                                   s.uw.copy(), s.vw.copy(), s.ww.copy()), tol=1e-12, max_iter=20000)
    return s, ref


@pytest.mark.parametrize("precision,tol", [(elq.FP64, 2e-6), (elq.MIXED, 2e-5)])
def test_128_against_oracle(ctx, drop128, precision, tol):
    s, ref = drop128
    p, idx, st = ctx.solve_pressure(s.copy(), elq.default_opts(tol=1e-8, max_iter=20000, precision=precision))
    assert st.true_rel_residual <= 1e-8
    np.testing.assert_array_equal(idx, ref.indices)
    assert np.abs(p - ref.pressure).max() <= tol * np.abs(ref.pressure).max()


def max_fluid_divergence(s, u, v, w, fluid=None, slab=32):
    """Largest |weighted divergence| (over fluid cells if given), in k slabs."""
    ni, nj, nk = s.ni, s.nj, s.nk
    uu, vv, ww = u.reshape(nk, nj, ni + 1), v.reshape(nk, nj + 1, ni), w.reshape(nk + 1, nj, ni)
    uw, vw, wwt = s.uw.reshape(nk, nj, ni + 1), s.vw.reshape(nk, nj + 1, ni), s.ww.reshape(nk + 1, nj, ni)
    best = 0.0
    for k0 in range(0, nk, slab):
        k1 = min(nk, k0 + slab)
        a = (uu[k0:k1] * uw[k0:k1]).astype(np.float64)
        b = (vv[k0:k1] * vw[k0:k1]).astype(np.float64)
        c = (ww[k0:k1 + 1] * wwt[k0:k1 + 1]).astype(np.float64)
        d = np.abs(((a[:, :, 1:] - a[:, :, :-1]) + (b[:, 1:, :] - b[:, :-1, :]) + (c[1:] - c[:-1])) / s.dx)
        if fluid is not None:
            d = d[fluid.reshape(nk, nj, ni)[k0:k1]]
        if d.size:
            best = max(best, float(d.max()))
    return best


def test_512_mixed_and_fp64():
    s = elq.scene_spheres(512, seed=1, n_spheres=128)
    div0 = max_fluid_divergence(s, s.u, s.v, s.w)
    c = elq.Context(0)
    ds = elq.DeviceScene(c, s)
    try:
        out = {}
        for prec in (elq.MIXED, elq.FP64):
            print(f"512^3 precision={prec}", flush=True)
            ds.reset_velocity()
            rc, st = c.project_device(ds.grid(), s.dt, elq.default_opts(tol=1e-8, max_iter=1000, precision=prec,
                                                                        check_every=32))
            assert rc == elq.OK and st.converged, (rc, c.last_error(), st.as_dict())
            assert st.true_rel_residual <= 1e-8, st.true_rel_residual
            assert 40_000_000 < st.n_unknowns < 2 ** 31 and 300 < st.iterations <= 1000, st.as_dict()
            idx = ds.fetch("indices", np.int32)
            fluid = idx != elq.UNKNOWN_INDEX
            assert int(fluid.sum()) == st.n_unknowns and int(idx[fluid].max()) == st.n_unknowns - 1
            div = max_fluid_divergence(s, ds.fetch("u"), ds.fetch("v"), ds.fetch("w"), fluid)
            assert div <= 1e-4 * div0, div / div0
            out[prec] = ds.fetch("pressure")
        scale = np.abs(out[elq.FP64]).max()
        assert np.abs(out[elq.MIXED] - out[elq.FP64]).max() <= 1e-5 * scale
    finally:
        ds.close()
        c.close()
