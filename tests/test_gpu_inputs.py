"""The frame's input steps before the projection on the GPU, through the C ABI,
against the oracle, bit for bit:

- elq_compute_weights_device: compute_staggered_grid_weight (boundary.cpp:6-34),
  every corner-sign pattern of the 4-corner fraction (level_set.h:83-163);
- elq_add_gravity_device: add_gravity (simulator.cpp:90-93);
- elq_valid_and_zero_device: find_valid_velocity + set_invalid_velocity_zero
  (simulator.cpp:125-178).
"""
import numpy as np
import pytest

import elq
import oracle as O

pytestmark = pytest.mark.gpu


def upload(ctx, a):
    p = ctx.alloc(max(a.nbytes, 4))
    ctx.h2d(p, a)
    return p


def download(ctx, p, like):
    out = np.empty_like(like)
    ctx.d2h(out, p)
    return out


def face_sizes(ni, nj, nk):
    return (ni + 1) * nj * nk, ni * (nj + 1) * nk, ni * nj * (nk + 1)


def nodal_room(ni, nj, nk):
    a = np.empty((ni + 1) * (nj + 1) * (nk + 1), np.float32)
    O.lib().orc_room_nodal(ni, nj, nk, a)
    return a


def nodal_random(ni, nj, nk, seed):
    rng = np.random.default_rng(seed)
    a = rng.uniform(-1.0, 1.0, (ni + 1) * (nj + 1) * (nk + 1)).astype(np.float32)
    a[rng.random(a.size) < 0.1] = 0.0       # sign ties: 0 counts as outside
    a[rng.random(a.size) < 0.02] = -0.0
    return a


def nodal_sphere(ni, nj, nk):
    z, y, x = np.meshgrid(np.arange(nk + 1), np.arange(nj + 1), np.arange(ni + 1), indexing="ij")
    d = np.sqrt((x - ni / 2) ** 2 + (y - nj / 2) ** 2 + (z - nk / 2) ** 2) - 0.35 * min(ni, nj, nk)
    return (-d).astype(np.float32).reshape(-1)   # solid outside a ball


CASES = {
    "room_20x14x9": lambda: (20, 14, 9, nodal_room(20, 14, 9)),
    "random_13x11x17": lambda: (13, 11, 17, nodal_random(13, 11, 17, 3)),
    "random_8x8x8": lambda: (8, 8, 8, nodal_random(8, 8, 8, 4)),
    "ball_24x20x16": lambda: (24, 20, 16, nodal_sphere(24, 20, 16)),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_weights_bit_exact(ctx, name):
    ni, nj, nk, nodal = CASES[name]()
    nu, nv, nw = face_sizes(ni, nj, nk)
    ref = [np.empty(n, np.float32) for n in (nu, nv, nw)]
    O.lib().orc_compute_weights(ni, nj, nk, nodal, *ref)
    pn = upload(ctx, nodal)
    pw = [ctx.alloc(n * 4) for n in (nu, nv, nw)]
    try:
        assert ctx.L.elq_compute_weights_device(ctx.h, ni, nj, nk, pn, *pw) == elq.OK
        for k, (p, r) in enumerate(zip(pw, ref)):
            np.testing.assert_array_equal(download(ctx, p, r).view(np.uint32), r.view(np.uint32), err_msg="uvw"[k])
    finally:
        for p in [pn] + pw:
            ctx.free(p)


def test_gravity_bit_exact(ctx):
    ni, nj, nk = 11, 9, 7
    rng = np.random.default_rng(5)
    v = rng.uniform(-3, 3, ni * (nj + 1) * nk).astype(np.float32)
    ref = v.copy()
    O.lib().orc_add_gravity(ni, nj, nk, ref, np.float32(0.01))
    pv = upload(ctx, v)
    try:
        assert ctx.L.elq_add_gravity_device(ctx.h, ni, nj, nk, pv, np.float32(0.01)) == elq.OK
        np.testing.assert_array_equal(download(ctx, pv, v).view(np.uint32), ref.view(np.uint32))
    finally:
        ctx.free(pv)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_valid_and_zero_bit_exact(ctx, seed):
    rng = np.random.default_rng(seed)
    s = elq.scene_spheres(20, seed=seed, n_spheres=5)
    ni, nj, nk = s.ni, s.nj, s.nk
    if seed == 3:   # random weights with zeros and a random level set with ties
        for a in (s.uw, s.vw, s.ww):
            a[:] = rng.uniform(0, 1, a.size).astype(np.float32)
            a[rng.random(a.size) < 0.2] = 0.0
        s.phi[:] = rng.uniform(-1, 1, s.phi.size).astype(np.float32)
        s.phi[rng.random(s.phi.size) < 0.1] = 0.0
    for a in (s.u, s.v, s.w):
        a[:] = rng.uniform(-1, 1, a.size).astype(np.float32)
    ref = [a.copy() for a in (s.u, s.v, s.w)]
    O.lib().orc_find_valid_and_zero(ni, nj, nk, s.phi, s.uw, s.vw, s.ww, *ref)
    ptr = {k: upload(ctx, getattr(s, k)) for k in ("phi", "u", "v", "w", "uw", "vw", "ww")}
    try:
        g = elq.Grid()
        g.ni, g.nj, g.nk, g.dx = ni, nj, nk, s.dx
        g.level_set, g.u, g.v, g.w = ptr["phi"], ptr["u"], ptr["v"], ptr["w"]
        g.u_weight, g.v_weight, g.w_weight = ptr["uw"], ptr["vw"], ptr["ww"]
        assert ctx.L.elq_valid_and_zero_device(ctx.h, g) == elq.OK
        for k, r in zip("uvw", ref):
            np.testing.assert_array_equal(download(ctx, ptr[k], r).view(np.uint32), r.view(np.uint32), err_msg=k)
        assert any((r == 0).any() for r in ref) and any((r != 0).any() for r in ref)
    finally:
        for p in ptr.values():
            ctx.free(p)
