"""The allocation-free projection (elq_project_device_noalloc) against the
projection it mirrors (elq_project_device), through the C ABI.

Both run the same kernels with the same launch configuration on the same
context, so every output is compared bit for bit: pressure, indices, the
updated u, v, w and the solver statistics. Two calls in a row, the second on
a smaller grid (the scratch array keeps the first call's contents; the
second must not see them). The context's own workspace stays empty across
the allocation-free calls (nothing is allocated for them), and grids whose
workspace does not fit the static array are refused with ELQ_ERR_ADMISSION:
up front when the dense arrays alone are too large, else after the index
pass, before the system is assembled or any output but the indices is written.
"""
import numpy as np
import pytest

import elq

pytestmark = pytest.mark.gpu

OUT = ("pressure", "indices", "u", "v", "w")


def _run(ctx, s, opts, noalloc):
    d = elq.DeviceScene(ctx, s)
    try:
        fn = ctx.project_device_noalloc if noalloc else ctx.project_device
        rc, st = fn(d.grid(), s.dt, opts)
        out = {k: d.fetch(k, np.int32 if k == "indices" else np.float32) for k in OUT}
        return rc, st, out
    finally:
        d.close()


def _same(a, b):
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("precision", [elq.MIXED, elq.FP64])
def test_noalloc_matches_project_device_twice(precision):
    print(f"noalloc precision={precision}", flush=True)
    opts = elq.default_opts(tol=1e-8, max_iter=5000, precision=precision)
    scenes = [elq.scene_spheres(64, seed=3, n_spheres=12), elq.scene_drop(32)]   # second call smaller
    ctx = elq.Context(0)
    try:
        got = []
        for s in scenes:
            rc, st, out = _run(ctx, s, opts, noalloc=True)
            assert rc == elq.OK and st.converged, (rc, ctx.last_error(), st.as_dict())
            got.append((st, out))
        assert ctx.device_bytes() == 0        # the context's workspace was never allocated
        for s, (st, out) in zip(scenes, got):
            rc, st_ref, ref = _run(ctx, s, opts, noalloc=False)
            assert rc == elq.OK
            for k in OUT:
                assert _same(out[k], ref[k]), k
            for f in ("n_unknowns", "n_bricks", "iterations", "outer_iterations", "error", "true_rel_residual",
                      "status", "converged"):
                assert getattr(st, f) == getattr(st_ref, f), f
        assert ctx.device_bytes() > 0         # while the allocating entry point did allocate
        # and the allocation-free call still agrees after the allocating ones ran
        rc, st, out = _run(ctx, scenes[1], opts, noalloc=True)
        assert rc == elq.OK and all(_same(out[k], got[1][1][k]) for k in OUT)
    finally:
        ctx.close()


def test_noalloc_refuses_grids_beyond_the_scratch_array():
    ctx = elq.Context(0)
    try:
        # dense arrays alone exceed 64 MiB: refused before any launch
        s = elq.make_scene(256, 256, 256, 0.3, [])
        d = elq.DeviceScene(ctx, s)
        try:
            ctx.h2d(d.ptr["indices"], np.full(s.cells, 7, np.int32))
            rc, st = ctx.project_device_noalloc(d.grid(), s.dt, elq.default_opts(tol=1e-8))
            assert rc == elq.ADMISSION and "scratch" in ctx.last_error()
            assert (d.fetch("indices", np.int32) == 7).all()
        finally:
            d.close()
        # dense arrays fit, the system of 2 M unknowns does not: refused after the index pass
        s = elq.scene_drop(128)
        u0 = s.u.copy()
        d = elq.DeviceScene(ctx, s)
        try:
            rc, st = ctx.project_device_noalloc(d.grid(), s.dt, elq.default_opts(tol=1e-8, precision=elq.MIXED))
            assert rc == elq.ADMISSION and "scratch" in ctx.last_error(), (rc, ctx.last_error())
            assert _same(d.fetch("u"), u0)
        finally:
            d.close()
        assert ctx.device_bytes() == 0
        # the context still projects afterwards
        s = elq.scene_drop(32)
        rc, st, out = _run(ctx, s, elq.default_opts(tol=1e-8), noalloc=True)
        assert rc == elq.OK and st.converged
    finally:
        ctx.close()
