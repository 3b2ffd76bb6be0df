"""Both sweeps of an iteration in one launch (mixed precision): the same
operations and reductions as the two launches, so the results are the same
bit for bit; also with the automatic choice (fused up to 32768 bricks) and
against the oracle through the existing parity tests, which run it on their
small grids."""
import numpy as np
import pytest

import elq

pytestmark = pytest.mark.gpu


def _scenes():
    return {"drop32": elq.scene_drop(32), "spheres48": elq.scene_spheres(48, seed=3, n_spheres=8),
            "request": elq.scene_request(100001), "clamp16": elq.scene_clamp(16)}


@pytest.mark.parametrize("name", ["drop32", "spheres48", "request", "clamp16"])
@pytest.mark.parametrize("precond", [elq.IC0, elq.MIC0])
def test_fused_equals_split(ctx, name, precond):
    out = {}
    for sched in (elq.WAVEFRONT_SPLIT, elq.WAVEFRONT_FUSED, elq.WAVEFRONT):
        s = _scenes()[name]
        o = elq.default_opts(tol=1e-8, max_iter=5000, precision=elq.MIXED, precond=precond, schedule=sched)
        if precond == elq.MIC0:
            o.tau = 0.97
        p, idx, st = ctx.solve_pressure(s, o)
        assert st.converged and st.true_rel_residual <= 1e-8, (sched, st.as_dict())
        out[sched] = (p, st.iterations, st.outer_iterations)
    ref = out[elq.WAVEFRONT_SPLIT]
    for sched in (elq.WAVEFRONT_FUSED, elq.WAVEFRONT):
        p, it, outer = out[sched]
        np.testing.assert_array_equal(p.view(np.uint32), ref[0].view(np.uint32))
        assert (it, outer) == (ref[1], ref[2])


def test_fp64_ignores_fusion(ctx):
    s = elq.scene_drop(32)
    res = []
    for sched in (elq.WAVEFRONT_SPLIT, elq.WAVEFRONT_FUSED):
        p, idx, st = ctx.solve_pressure(s.copy(), elq.default_opts(tol=1e-8, max_iter=5000, precision=elq.FP64,
                                                                   schedule=sched))
        res.append(p)
    np.testing.assert_array_equal(res[0].view(np.uint32), res[1].view(np.uint32))
