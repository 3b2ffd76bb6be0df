"""Writes the golden fixtures of the projection from the CPU oracle.

Usage: python tests/golden/make_golden.py   (run from the repository root)

Each fixture holds the inputs of one projection (reference layout, float32)
and the oracle's outputs: indices, the assembled system in natural stencil
form (diag, ax, ay, az, b), IC(0) pivots, the pressure at tol 1e-12 and the
updated face velocities. The scenes are the drop scene (slab y < 0.4 plus a
sphere) and a seeded slab + spheres scene, small enough to keep each file a
few hundred KB.
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


if __name__ == "__main__":
    for name, make in SCENES.items():
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **fixture(make()))
        print("wrote", path, os.path.getsize(path), "bytes")
