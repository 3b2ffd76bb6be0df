"""Batched serving (elq_server_opts.batched): the Krylov iterations of the
running requests advance together, one launch per kernel class for all of
them, replayed from a HIP graph captured once (elq_server.cpp, Batcher).

- A request's workgroups in a batched launch are those its own launches
  would use, so its outputs are bit-identical batched or not, and alone or
  in any mix.
- Requests really share launches: the mean batch size over a mix of ten is
  above one.
- Small requests served batched in a mix with larger ones match the
  oracle's FP64 solve of the same scene to rel.res 1e-12: the same unknown
  count, pressure within 2e-5 of max|p| (mixed precision).
- FP64 servers ignore the flag (the batched kernels are FP32): no batches.
"""
import numpy as np
import pytest

import elq
from test_gpu_serving import TOL, opts, outputs, requests, run

pytestmark = pytest.mark.gpu


def _same(a, b):
    assert a["status"] == elq.OK and b["status"] == elq.OK, (a, b)
    assert a["iterations"] == b["iterations"] and a["outer_iterations"] == b["outer_iterations"]
    for x, y, name in zip(outputs(a), outputs(b), ("pressure", "u", "v", "w")):
        np.testing.assert_array_equal(x.view(np.uint32), y.view(np.uint32), err_msg=name)


def test_batched_equals_direct_and_shares_launches():
    seeds = list(range(100, 110))
    with elq.Server(0, max_slots=4, opts=opts(), batched=True) as srv:
        batched = run(srv, requests(seeds))
        nb, nj = srv.batch_stats()
    with elq.Server(0, max_slots=4, opts=opts()) as srv:
        direct = run(srv, requests(seeds))
        assert srv.batch_stats() == (0, 0)
    for sd in seeds:
        _same(batched[sd], direct[sd])
        assert batched[sd]["true_rel_residual"] <= TOL
    assert nb > 0 and nj / nb > 1.2, (nb, nj)


def test_batched_alone_equals_mixed():
    target = 104
    with elq.Server(0, max_slots=4, opts=opts(), batched=True) as srv:
        alone = run(srv, {target: elq.scene_request(target)})[target]
        mixed = run(srv, requests(range(100, 110)))
    _same(alone, mixed[target])


def test_batched_small_requests_match_oracle():
    import oracle as O
    seeds = [418, 429, 437]
    scenes = requests(seeds)
    mix = {k: v.copy() for k, v in scenes.items()}
    mix.update(requests([104, 105]))
    with elq.Server(0, max_slots=3, opts=opts(), batched=True) as srv:
        got = run(srv, mix)
    for sd in seeds:
        r, sc = got[sd], scenes[sd]
        assert r["status"] == elq.OK and r["true_rel_residual"] <= TOL, r
        ref = O.solve_pressure(O.Scene(sc.ni, sc.nj, sc.nk, sc.dx, sc.dt, sc.phi.copy(), sc.u.copy(), sc.v.copy(),
                                       sc.w.copy(), sc.uw.copy(), sc.vw.copy(), sc.ww.copy()), tol=1e-12,
                               max_iter=20000)
        assert r["n_unknowns"] == ref.n
        assert np.abs(r["pressure"] - ref.pressure).max() <= 2e-5 * np.abs(ref.pressure).max()


def test_fp64_server_ignores_batching():
    with elq.Server(0, max_slots=2, opts=opts(elq.FP64), batched=True) as srv:
        got = run(srv, requests([418, 429]))
        assert srv.batch_stats() == (0, 0)
    assert all(r["status"] == elq.OK and r["true_rel_residual"] <= TOL for r in got.values())
