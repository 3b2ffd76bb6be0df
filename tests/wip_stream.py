"""Stream sweeps (schedule ELQ_SCHED_STREAM, mixed precision) against the
CPU oracle, through the C ABI: the same scenes and bounds as the wavefront
schedule in test_gpu_parity.py (pressure within 2e-5 of max|p| of the
oracle's 1e-12 solve, true FP64 relative residual <= 1e-8), plus grids whose
blocks outnumber the workgroups (several rounds per workgroup), axes of
4092 cells and a solve stopped by max_iter, whose iterate and error match
the wavefront schedule's."""
import numpy as np
import pytest

import elq
import oracle as O
from test_gpu_parity import SCENES, _check_solution

pytestmark = pytest.mark.gpu


def _opts(**kw):
    base = dict(tol=1e-8, max_iter=5000, precision=elq.MIXED, schedule=elq.STREAM)
    base.update(kw)
    return elq.default_opts(**base)


@pytest.mark.parametrize("name", ["drop16", "drop32", "spheres24", "aniso", "odd", "clamp16"])
def test_stream_projection(ctx, name):
    s = SCENES[name]()
    g = s.copy()
    p, idx, st = ctx.solve_pressure(g, _opts())
    assert st.status == elq.OK and st.converged, st.as_dict()
    ref = _check_solution(s, p, idx, st, 2e-5)
    umax = max(np.abs(ref.u).max(), np.abs(ref.v).max(), np.abs(ref.w).max())
    for f in ("u", "v", "w"):
        assert np.abs(getattr(g, f) - getattr(ref, f)).max() <= 2e-5 * umax, f


@pytest.mark.parametrize("name", ["long_x", "long_y", "long_z"])
def test_stream_long_axis(ctx, name):
    s = SCENES[name]()
    p, idx, st = ctx.solve_pressure(s.copy(), _opts())
    assert st.status == elq.OK and st.converged, st.as_dict()
    _check_solution(s, p, idx, st, 2e-5)


@pytest.mark.parametrize("cap", [1, 3, 8])
def test_stream_rounds(cap):
    """Fewer workgroups than blocks: each workgroup sweeps several blocks in
    list order (a single workgroup: all of them)."""
    c = elq.Context(0)
    try:
        c.set_max_groups(cap)
        s = SCENES["spheres24"]()
        p, idx, st = c.solve_pressure(s.copy(), _opts())
        assert st.status == elq.OK and st.converged, st.as_dict()
        _check_solution(s, p, idx, st, 2e-5)
    finally:
        c.close()


def test_stream_iterations_match_wavefront(ctx):
    """Same algebra as the wavefront sweeps: the iteration counts agree to
    within a few (the sums of a sweep step run in another order), and a
    solve stopped by max_iter reports the same count."""
    s = elq.scene_drop(48)
    _, _, a = ctx.solve_pressure(s.copy(), _opts(schedule=elq.WAVEFRONT))
    _, _, b = ctx.solve_pressure(s.copy(), _opts())
    assert a.converged and b.converged
    assert abs(a.iterations - b.iterations) <= 3, (a.iterations, b.iterations)
    _, _, c = ctx.solve_pressure(s.copy(), _opts(max_iter=7, check_every=1))
    assert c.status == elq.NOT_CONVERGED and c.iterations == 7, c.as_dict()
