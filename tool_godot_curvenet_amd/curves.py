"""Closed cubic B-spline curves driving the constrained (seam) vertices.

The curve -> constraint step in front of the harmonic solve: each cut loop
carries a closed uniform cubic B-spline with `n_ctrl` control values (per
coordinate), and every constrained vertex takes the spline's value at its
arc-length parameter on the loop.  The map control values -> constrained
values is linear, a sparse (n_con x n_loops * n_ctrl) matrix with 4 entries
per row, so the adjoint's dL/dx_c chains to the control points by its
transpose.  (The reference samples its curves with Bezier/B-spline
evaluation in EndToEndExample.lean:52-60 before projecting knots to
vertices, SurfaceProjection.lean:46-80; here the vertices are fixed by the
cut and the spline is evaluated at them.)
"""
from __future__ import annotations

import numpy as np


def bspline_weights(t: np.ndarray, n_ctrl: int):
    """Closed uniform cubic B-spline with n_ctrl control points, at
    parameters t in [0, 1).  Returns (idx, w), both (len(t), 4): the value is
    sum_q w[:, q] * ctrl[idx[:, q]]."""
    u = np.mod(np.asarray(t, dtype=np.float64), 1.0) * n_ctrl
    i = np.floor(u).astype(np.int64)
    f = u - i
    f2, f3 = f * f, f * f * f
    w = np.stack([(1.0 - f) ** 3 / 6.0,
                  (3.0 * f3 - 6.0 * f2 + 4.0) / 6.0,
                  (-3.0 * f3 + 3.0 * f2 + 3.0 * f + 1.0) / 6.0,
                  f3 / 6.0], axis=1)
    idx = np.stack([(i - 1) % n_ctrl, i % n_ctrl, (i + 1) % n_ctrl, (i + 2) % n_ctrl], axis=1)
    return idx, w


def loop_parameters(pos: np.ndarray, loop) -> np.ndarray:
    """Arc-length parameter in [0, 1) of every vertex of a closed loop."""
    p = pos[np.asarray(loop)]
    seg = np.linalg.norm(np.roll(p, -1, axis=0) - p, axis=1)
    s = np.concatenate([[0.0], np.cumsum(seg)[:-1]])
    return s / seg.sum()


def constrained_basis_from(origin, con_of, n_con: int, mesh, n_ctrl: int = 16) -> np.ndarray:
    """Dense (n_con x n_loops * n_ctrl) basis: row c holds the spline weights
    of constrained vertex c (`con_of[v] = c` for cut-mesh vertex v, `origin[v]`
    its mesh vertex).  A vertex (or its seam copy) takes the parameter of its
    origin on the first loop that contains it."""
    loops = mesh.meta["loop_verts"]
    where = {}
    for L, loop in enumerate(loops):
        t = loop_parameters(mesh.pos, loop)
        for v, tv in zip(loop, t):
            where.setdefault(int(v), (L, float(tv)))
    origin = np.asarray(origin)
    con_of = np.asarray(con_of)
    Bm = np.zeros((n_con, len(loops) * n_ctrl))
    for v in np.nonzero(con_of >= 0)[0]:
        L, tv = where[int(origin[v])]
        idx, w = bspline_weights(np.array([tv]), n_ctrl)
        np.add.at(Bm[con_of[v]], L * n_ctrl + idx[0], w[0])
    return Bm


def constrained_basis(lap, mesh, n_ctrl: int = 16) -> np.ndarray:
    """constrained_basis_from for a CutLaplacian built by from_mesh."""
    _, con_of = lap.maps()
    return constrained_basis_from(lap.origin, con_of, lap.n_con, mesh, n_ctrl)
