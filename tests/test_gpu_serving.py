"""Serving (continuous batching) through the C ABI: elq_server_*.

- A request's outputs are bit-identical alone and in a mix: every slot runs
  the same kernels with the same grid caps, so nothing of the arithmetic
  depends on which other requests share the device.
- Every result matches a single-context projection of the same scene to the
  solve tolerance (the slot's grid cap changes only the reduction order):
  max|p - p_single| <= 2e-5 max|p_single| (mixed) and true FP64 relative
  residual <= 1e-8.
- Small requests (every axis <= 48) served in a mix match the oracle's FP64
  solve of the same scene to rel.res 1e-12: the same unknown count, pressure
  within 2e-6 (FP64) / 2e-5 (mixed) of max|p|, velocities within that bound
  times dt / (0.01 dx), the largest factor the update applies.
- Admission never lets the slots together hold more device memory than the
  budget; a request that alone exceeds it is refused at submit.
- The footprint model is an upper bound of what a projection allocates.
"""
import numpy as np
import pytest

import elq

pytestmark = pytest.mark.gpu

TOL = 1e-8


def opts(precision=elq.MIXED):
    return elq.default_opts(tol=TOL, max_iter=20000, precision=precision, check_every=16)


def requests(seeds):
    return {sd: elq.scene_request(sd) for sd in seeds}


def outputs(r):
    s = r["scene"]
    return r["pressure"], s.u, s.v, s.w


def run(server, scenes):
    ids = {server.submit(sc): key for key, sc in scenes.items()}
    got = {}
    while len(got) < len(ids):
        assert server.wait(120000) > 0, "no request finished within 120 s"
        for r in server.poll():
            got[ids[r["id"]]] = r
    return got


@pytest.mark.parametrize("precision", [elq.FP64, elq.MIXED])
def test_footprint_bounds_allocation(precision):
    for s in (elq.scene_drop(32), elq.scene_request(5), elq.scene_spheres(40, seed=2, n_spheres=5)):
        with elq.Context(0) as c:
            c.solve_pressure(s.copy(), opts(precision))
            held = c.device_bytes()
        assert 0 < held <= elq.footprint(s, precision), (held, elq.footprint(s, precision))
        assert elq.footprint(s, precision) <= elq.footprint(s, precision, bound_by_level_set=False)


def test_alone_equals_mixed():
    seeds = list(range(100, 110))
    target = 104
    with elq.Server(0, max_slots=4, opts=opts()) as srv:
        alone = run(srv, {target: elq.scene_request(target)})[target]
        mixed = run(srv, requests(seeds))
    m = mixed[target]
    assert alone["status"] == elq.OK and m["status"] == elq.OK
    assert alone["iterations"] == m["iterations"] and alone["outer_iterations"] == m["outer_iterations"]
    for a, b, name in zip(outputs(alone), outputs(m), ("pressure", "u", "v", "w")):
        np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32), err_msg=name)
    # the mix really overlapped on the device
    spans = sorted((r["t_admit_ms"], r["t_done_ms"]) for r in mixed.values())
    assert any(spans[i + 1][0] < spans[i][1] for i in range(len(spans) - 1))


def test_results_match_single_context(ctx):
    seeds = [200, 201, 202, 203, 204, 205]
    scenes = requests(seeds)
    with elq.Server(0, max_slots=3, opts=opts()) as srv:
        got = run(srv, {k: v.copy() for k, v in scenes.items()})
    for sd in seeds:
        r = got[sd]
        assert r["status"] == elq.OK, r
        assert r["true_rel_residual"] <= TOL
        p_ref, _, st = ctx.solve_pressure(scenes[sd].copy(), opts())
        assert r["n_unknowns"] == st.n_unknowns
        scale = np.abs(p_ref).max()
        assert np.abs(r["pressure"] - p_ref).max() <= 2e-5 * scale


@pytest.mark.parametrize("precision, tol", [(elq.FP64, 2e-6), (elq.MIXED, 2e-5)])
def test_small_requests_match_oracle(precision, tol):
    import oracle as O
    seeds = [418, 429, 437]
    scenes = requests(seeds)
    assert all(max(sc.ni, sc.nj, sc.nk) <= 48 for sc in scenes.values())
    mix = {k: v.copy() for k, v in scenes.items()}
    mix.update({k: v for k, v in requests([104, 105]).items()})   # larger requests in the same mix
    with elq.Server(0, max_slots=3, opts=opts(precision)) as srv:
        got = run(srv, mix)
    for sd in seeds:
        r, sc = got[sd], scenes[sd]
        assert r["status"] == elq.OK and r["true_rel_residual"] <= TOL, r
        ref = O.solve_pressure(O.Scene(sc.ni, sc.nj, sc.nk, sc.dx, sc.dt, sc.phi.copy(), sc.u.copy(), sc.v.copy(),
                                       sc.w.copy(), sc.uw.copy(), sc.vw.copy(), sc.ww.copy()), tol=1e-12,
                               max_iter=20000)
        assert r["n_unknowns"] == ref.n
        scale = np.abs(ref.pressure).max()
        assert np.abs(r["pressure"] - ref.pressure).max() <= tol * scale
        served = r["scene"]
        for f in ("u", "v", "w"):
            bound = tol * scale * sc.dt / (sc.dx * 0.01) + 1e-6 * np.abs(getattr(ref, f)).max()
            assert np.abs(getattr(served, f) - getattr(ref, f)).max() <= bound, f


def test_admission_respects_budget():
    scenes = requests(range(300, 312))
    fps = {k: elq.footprint(s, elq.MIXED) for k, s in scenes.items()}
    budget = int(max(fps.values()) * 1.6)
    big = elq.scene_spheres(128, seed=3, n_spheres=4)
    with elq.Server(0, max_slots=4, hbm_budget=budget, opts=opts()) as srv:
        if elq.footprint(big, elq.MIXED) > budget:
            with pytest.raises(elq.ElqError) as e:
                srv.submit(big)
            assert e.value.code == elq.ADMISSION
        got = run(srv, scenes)
        assert srv.device_bytes() <= budget
    assert all(r["status"] == elq.OK for r in got.values())
    # at any instant the footprints of the running requests fit the budget
    events = sorted([(r["t_admit_ms"], fps[k]) for k, r in got.items()] +
                    [(r["t_done_ms"], -fps[k]) for k, r in got.items()], key=lambda e: (e[0], e[1]))
    live = 0
    for _, d in events:
        live += d
        assert live <= budget


def test_queue_full_reports_busy():
    scenes = [elq.scene_spheres(64, seed=9, n_spheres=8) for _ in range(6)]
    with elq.Server(0, max_slots=1, max_queue=2, opts=opts()) as srv:
        codes = []
        for s in scenes:
            try:
                srv.submit(s)
                codes.append(elq.OK)
            except elq.ElqError as e:
                codes.append(e.code)
        assert elq.BUSY in codes
        while srv.in_flight():
            srv.wait(60000)
            srv.poll()
