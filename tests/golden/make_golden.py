"""Writes the golden fixtures of the projection from the CPU oracle.

Usage: python tests/golden/make_golden.py [name ...]   (run from the repository root)

Each projection fixture holds the inputs of one projection (reference
layout, float32) and the oracle's outputs: indices, the assembled system in
natural stencil form (diag, ax, ay, az, b), IC(0) pivots, the pressure at tol
1e-12 and the updated face velocities. The scenes are the drop scene (slab
y < 0.4 plus a sphere) and a seeded slab + spheres scene, small enough to
keep each file a few hundred KB.

The frame fixture (frame16.npz) continues the seeded scene after its
projection through the reference's next frame steps: the face flags
(find_valid_velocity), extrapolate_velocity (10 layers), advect_level_set,
advect_velocity, and the reinitialisation of the advected level set with the
room's level-set flags.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
SCENES = {
    "drop12": lambda: O.scene_drop(12),
    "spheres16": lambda: O.scene_spheres(16, seed=21, n_spheres=5),
    # interface 1e-5 from cell centres on every axis: ghost-fluid faces with
    # theta < 0.01, i.e. the clamp of projection.cpp:59-63,100-131
    "clamp16": lambda: O.scene_clamp(16),
}


def fixture(s):
    n, idx = O.assign_indices(s)
    sys_ = O.build_matrix(s, idx, n)
    inv_e, _ = O.ic_factor(sys_)
    res = O.solve_pressure(s, tol=1e-12, max_iter=10000)
    return dict(ni=s.ni, nj=s.nj, nk=s.nk, dx=np.float32(s.dx), dt=np.float32(s.dt), phi=s.phi, u=s.u, v=s.v,
                w=s.w, uw=s.uw, vw=s.vw, ww=s.ww, n=n, indices=idx, diag=sys_.diag, ax=sys_.ax, ay=sys_.ay,
                az=sys_.az, b=sys_.b, inv_e=inv_e, pressure=res.pressure, u_out=res.u, v_out=res.v,
                w_out=res.w, iterations=res.iterations)


def frame_fixture():
    s = O.scene_spheres(16, seed=21, n_spheres=5)
    res = O.solve_pressure(s, tol=1e-12, max_iter=10000)
    s.u, s.v, s.w = res.u.copy(), res.v.copy(), res.w.copy()
    ni, nj, nk = s.ni, s.nj, s.nk
    flags = [f.copy() for f in O.find_valid_velocity(s)]
    ext = [a.copy() for a in (s.u, s.v, s.w)]
    ext_flags = [f.copy() for f in flags]
    for a, f, d in zip(ext, ext_flags, ((ni + 1, nj, nk), (ni, nj + 1, nk), (ni, nj, nk + 1))):
        O.extrapolate(a, f, d, 10)
    adv = O.Scene(ni, nj, nk, s.dx, s.dt, s.phi.copy(), ext[0].copy(), ext[1].copy(), ext[2].copy(),
                  s.uw, s.vw, s.ww)
    O.advect_level_set(adv, s.dt)
    O.advect_velocity(adv, s.dt)
    nodal = np.empty((ni + 1) * (nj + 1) * (nk + 1), np.float32)
    O.lib().orc_room_nodal(ni, nj, nk, nodal)
    ls_flags = np.empty(ni * nj * nk, np.uint8)
    O.lib().orc_find_valid_level_set(ni, nj, nk, nodal, ls_flags)
    rei = adv.phi.copy()
    left, passes = O.reinitialize(rei, ls_flags, (ni, nj, nk), s.dx)
    return dict(ni=ni, nj=nj, nk=nk, dx=np.float32(s.dx), dt=np.float32(s.dt), phi=s.phi, u=s.u, v=s.v, w=s.w,
                uw=s.uw, vw=s.vw, ww=s.ww, u_flags=flags[0], v_flags=flags[1], w_flags=flags[2],
                u_ext=ext[0], v_ext=ext[1], w_ext=ext[2], u_ext_flags=ext_flags[0], v_ext_flags=ext_flags[1],
                w_ext_flags=ext_flags[2], phi_adv=adv.phi, u_adv=adv.u, v_adv=adv.v, w_adv=adv.w,
                ls_flags=ls_flags, phi_rei=rei, rei_passes=passes, rei_unreached=left)


if __name__ == "__main__":
    # optional arguments: the fixture names to (re)write; default all
    want = set(sys.argv[1:])
    for name, make in SCENES.items():
        if want and name not in want:
            continue
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **fixture(make()))
        print("wrote", path, os.path.getsize(path), "bytes")
    if want and "frame16" not in want:
        sys.exit(0)
    path = os.path.join(HERE, "frame16.npz")
    np.savez_compressed(path, **frame_fixture())
    print("wrote", path, os.path.getsize(path), "bytes")
