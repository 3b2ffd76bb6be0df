"""Uniform interior bricks against the oracle.

The SpMV and the wavefront sweeps read the coefficients of a brick whose 64
cells all have couplings -1 and diagonal 6 (FP32: ghost part 0) from a
64-entry constant row, and in FP64 also its pivot inverses when they all
equal the interior fixed point of the factorisation. Such a brick must also
have -1 couplings on its -x/-y/-z faces, which the neighbour bricks store.

The scene is a box of fluid with one row of cells, (i, j) = (16, 12) for k
in [8, 24), that only those face couplings tell apart: i = 16 is the first
cell layer of brick column bx = 4 and j = 12 the first of by = 3. Their -x
face has weight 0.5 (coupling -0.5 to a fluid cell, stored in brick
bx = 3) and their -y neighbour is an air cell with level set exactly 0
(theta = 1) behind a face of weight 1.5, so their FP64 diagonal is still
0.5 + 1.5 + 4 = 6 and their own (+x/+y/+z) couplings are -1. Flagging the
brick would make the FP64 SpMV use -1 instead of -0.5 against a nonzero
neighbour value. (In FP32 the ghost part 1.5 already keeps it unflagged.)
The solves are compared with the oracle's (IC(0) and MIC(0)).
"""
import numpy as np
import pytest

import elq
import oracle as O

pytestmark = pytest.mark.gpu

N = 32
I0, J0 = 16, 12
KS = slice(8, 24)


def _box_with_odd_row():
    s = elq.scene_drop(N)
    phi = s.phi.reshape(N, N, N)  # [k][j][i]
    phi[:] = -0.5
    phi[KS, J0 - 1, I0] = 0.0  # air cells below the row, theta = 1 for it
    s.uw.reshape(N, N, N + 1)[KS, J0, I0] = 0.5  # -x faces of the row
    s.vw.reshape(N, N + 1, N)[KS, J0, I0] = 1.5  # -y faces (to the air cells)
    return s


def _to_oracle(s):
    return O.Scene(s.ni, s.nj, s.nk, s.dx, s.dt, s.phi.copy(), s.u.copy(), s.v.copy(), s.w.copy(),
                   s.uw.copy(), s.vw.copy(), s.ww.copy())


def test_scene_has_a_brick_only_its_face_couplings_tell_apart():
    s = _box_with_odd_row()
    os_ = _to_oracle(s)
    n, idx = O.assign_indices(os_)
    sys_ = O.build_matrix(os_, idx, n)
    I = idx.reshape(N, N, N)
    row = I[KS, J0, I0]
    assert ((row >= 0) & (row < n)).all()
    # own entries as in the interior ...
    assert (sys_.diag[row] == 6.0).all()
    for a in (sys_.ax, sys_.ay, sys_.az):
        assert (a[row] == -1.0).all()
    # ... but the -x coupling (stored with the fluid -x neighbour) is -0.5
    left = I[KS, J0, I0 - 1]
    assert ((left >= 0) & (left < n)).all() and (sys_.ax[left] == -0.5).all()
    # the other cells of those bricks are interior ones
    blk = I[8:24, 12:16, 16:20]
    assert (sys_.diag[blk] == 6.0).all() and (sys_.ax[blk] == -1.0).all() and (sys_.az[blk] == -1.0).all()


@pytest.mark.parametrize("precision, tol", [(elq.FP64, 2e-6), (elq.MIXED, 2e-5)])
@pytest.mark.parametrize("precond, tau", [(elq.IC0, 0.0), (elq.MIC0, 0.97)])
def test_solve_next_to_non_uniform_faces(ctx, precision, tol, precond, tau):
    s = _box_with_odd_row()
    g = s.copy()
    o = elq.default_opts(tol=1e-8, max_iter=5000, precision=precision, precond=precond)
    if precond == elq.MIC0:
        o.tau = tau
    p, idx, st = ctx.solve_pressure(g, o)
    assert st.converged and st.true_rel_residual <= 1e-8, st.as_dict()
    ref = O.solve_pressure(_to_oracle(s), tol=1e-12, max_iter=10000, tau=tau)
    np.testing.assert_array_equal(idx, ref.indices)
    assert np.abs(p - ref.pressure).max() <= tol * np.abs(ref.pressure).max()
