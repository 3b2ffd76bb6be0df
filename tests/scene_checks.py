"""Checks on the test scenes themselves (no oracle arithmetic restated)."""
import numpy as np

import oracle as O


def clamped_faces(s):
    """Faces between an unknown and a valid air neighbour whose ghost-fluid
    fraction falls below 0.01 (the clamp of projection.cpp:59-63,100-131), per
    axis (x, y, z). The fraction is taken in float64 from the level set, so
    the count does not depend on the oracle's float arithmetic."""
    ni, nj, nk = s.ni, s.nj, s.nk
    phi = s.phi.reshape(nk, nj, ni).astype(np.float64)
    n, idx = O.assign_indices(s)
    fluid = (idx != O.UNKNOWN_INDEX).reshape(nk, nj, ni)
    uw = s.uw.reshape(nk, nj, ni + 1)
    vw = s.vw.reshape(nk, nj + 1, ni)
    ww = s.ww.reshape(nk + 1, nj, ni)
    valid = (uw[:, :, :-1] > 0) | (uw[:, :, 1:] > 0) | (vw[:, :-1] > 0) | (vw[:, 1:] > 0) | \
            (ww[:-1] > 0) | (ww[1:] > 0)
    air = valid & (phi >= 0)
    out = []
    for ax in (2, 1, 0):
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[ax] = slice(0, -1)
        hi[ax] = slice(1, None)
        a, b = phi[tuple(lo)], phi[tuple(hi)]
        with np.errstate(divide="ignore", invalid="ignore"):
            th_ab = np.where(a < 0, a / (a - b), 1.0)
            th_ba = np.where(b < 0, b / (b - a), 1.0)
        cnt = int((fluid[tuple(lo)] & air[tuple(hi)] & (th_ab < 0.01)).sum() +
                  (fluid[tuple(hi)] & air[tuple(lo)] & (th_ba < 0.01)).sum())
        out.append(cnt)
    return out
