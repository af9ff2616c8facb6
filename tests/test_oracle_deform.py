"""Pins of the oracle's deformation pipeline (CPU only): the
DeformSolve.lean:164-253, EndToEndExample.lean and ScaledFrames.lean known
answers, the sparse stage-2 right-hand side against the dense spec
transcription, and the two-stage solve reproducing a rigid motion."""
import math

import numpy as np
import pytest

import oracle as O
from kat_data import EQ_TRI, TRIANGLE, he

R90 = np.array([[0.0, -1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 1.0]])
X, Y = np.array([1.0, 0.0, 0.0]), np.array([0.0, 1.0, 0.0])
O0 = np.zeros(3)


def test_scaled_frames_kats():                      # ScaledFrames.lean:156-260
    assert np.allclose(O.isolated_segment_gradient(O0, X, O0, X), np.eye(3), atol=1e-12)
    assert np.allclose(O.isolated_segment_gradient(O0, X, O0, Y), R90, atol=1e-12)
    assert np.allclose(O.isolated_segment_gradient(O0, X, O0, 2 * X), 2 * np.eye(3), atol=1e-12)
    F = O.isolated_segment_gradient(O0, X, O0, 3 * Y)
    assert np.allclose(F @ X, 3 * Y, atol=1e-12)
    assert abs(O.lib().orc_mat3_det(np.eye(3).ctypes.data_as(O.P)) - 1.0) < 1e-12
    D = np.diag([2.0, 3.0, 4.0])
    assert abs(O.lib().orc_mat3_det(np.ascontiguousarray(D).ctypes.data_as(O.P)) - 24.0) < 1e-12
    assert np.allclose(O.mat3_inv(D), np.diag([0.5, 1 / 3, 0.25]), atol=1e-12)
    M = np.array([[2.0, 1.0, 0.0], [1.0, 2.0, 1.0], [0.0, 1.0, 2.0]])
    assert np.allclose(O.mat3_inv(M) @ M, np.eye(3), atol=1e-10)
    assert np.allclose(O.deformation_gradient(np.eye(3), D), D, atol=1e-12)
    assert np.allclose(O.deformation_gradient(M, M), np.eye(3), atol=1e-10)


def test_rotation_matches_rodrigues_with_angle():
    """Independent check: the angle-free Rodrigues form equals the textbook
    axis-angle rotation taking one unit vector onto another."""
    rng = np.random.default_rng(1)
    for _ in range(50):
        a, b = rng.normal(size=(2, 3))
        a /= np.linalg.norm(a)
        b /= np.linalg.norm(b)
        F = O.isolated_segment_gradient(O0, a, O0, b)
        ax = np.cross(a, b)
        s, c = np.linalg.norm(ax), float(a @ b)
        k = ax / s
        K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
        R = np.eye(3) + s * K + (1 - c) * K @ K
        assert np.allclose(F, R, atol=1e-10)


@pytest.mark.parametrize("Fc,Xc,v1,v2", [
    (np.eye(3), [5.0, 0.0, 0.0], [6.0, 0.0, 0.0], [5.5, 0.0, 0.8660254037844386]),
    (np.eye(3), [0.0, 2.0, 0.0], [1.0, 2.0, 0.0], None),
    (R90, [0.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.5, 0.8660254037844386]),
    (2 * np.eye(3), [0.0, 0.0, 0.0], [2.0, 0.0, 0.0], [1.0, 0.0, 1.7320508075688772]),
])
def test_deform_solve_kats(Fc, Xc, v1, v2):          # DeformSolve.lean:164-253
    h = he(TRIANGLE)
    Xv = O.solve_deformation_dense(h, [1, 0, 0], EQ_TRI, np.reshape(Fc, (1, 9)),
                                   np.reshape(Xc, (1, 3)))
    assert np.allclose(Xv[1], v1, atol=1e-9)
    if v2 is not None:
        assert np.allclose(Xv[2], v2, atol=1e-9)
    assert np.all(Xv[0] == 0)


def test_end_to_end_bezier_translation():            # EndToEndExample.lean:62-211
    rest = np.array([[-1.5, 0, 0], [-0.5, 0, 0], [0.5, 0, 0], [1.5, 0, 0]])
    posed = rest + np.array([5.0, 0.0, 0.0])
    rp, rq = O.bezier3(rest, 0.5), O.bezier3(rest, 0.6)
    pp, pq = O.bezier3(posed, 0.5), O.bezier3(posed, 0.6)
    assert np.abs(rp).max() < 1e-12 and np.abs(pp - [5.0, 0, 0]).max() < 1e-12
    Fc = O.isolated_segment_gradient(rp, rq, pp, pq)
    assert np.abs(Fc - np.eye(3)).max() < 1e-12
    h = he(TRIANGLE)
    Xv = O.solve_deformation_dense(h, [1, 0, 0], EQ_TRI, Fc.reshape(1, 9), pp.reshape(1, 3))
    assert np.allclose(Xv[1], [6.0, 0.0, 0.0], atol=1e-9)
    assert np.allclose(Xv[2], [5.5, 0.0, 0.8660254037844386], atol=1e-9)
    # rotation case: the posed curve is the rest curve turned 90 degrees about z
    rot = rest @ R90.T
    qp, qq = O.bezier3(rot, 0.5), O.bezier3(rot, 0.6)
    Fr = O.isolated_segment_gradient(rp, rq, qp, qq)
    assert np.abs(Fr - R90).max() < 1e-12
    Xr = O.solve_deformation_dense(h, [1, 0, 0], EQ_TRI, Fr.reshape(1, 9), qp.reshape(1, 3))
    assert np.allclose(Xr[1], [0.0, 1.0, 0.0], atol=1e-9)
    assert np.allclose(Xr[2], [0.0, 0.5, 0.8660254037844386], atol=1e-9)


def _patch(seed, n=7):
    """A jittered n x n patch cut along its middle row (the patch of the
    VtLhV dense-vs-CSR tests)."""
    rng = np.random.default_rng(seed)
    tris = []
    for r in range(n - 1):
        for c in range(n - 1):
            a, b, d, e = r * n + c, r * n + c + 1, (r + 1) * n + c, (r + 1) * n + c + 1
            tris += [(a, b, e), (a, e, d)] if rng.random() < 0.5 else [(a, b, d), (b, e, d)]
    tris = np.asarray(tris, np.int32)
    y, x = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    pos = np.stack([x.ravel(), y.ravel(), rng.uniform(-0.3, 0.3, n * n)], 1) + \
        rng.uniform(-0.2, 0.2, (n * n, 3))
    cut = np.array([(3 * n + c, 3 * n + c + 1) for c in range(1, n - 2)], np.int32)
    nt, origin, con, nv2 = O.cut_edges(n * n, tris, cut)
    # the open boundary is constrained too, so every free row is interior
    h = O.he_from_triangles(nv2, nt)
    bnd = np.unique(np.asarray(h["target"])[np.asarray(h["face"]) < 0])
    con = np.asarray(con, np.uint8).copy()
    con[bnd] = 1
    return h, con, pos[origin]


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_deform_rhs_csr_equals_dense(seed):
    h, con, pos = _patch(seed)
    r = O.cut_laplacian_csr(h, con, pos)
    free_of, con_of = r["free_of"], r["con_of"]
    nf, nc = r["n_free"], int(con.sum())
    rng = np.random.default_rng(seed)
    P = 2
    Ffree, Fc = rng.normal(size=(nf, 9 * P)), rng.normal(size=(nc, 9 * P))
    Xc = rng.normal(size=(nc, 3 * P))
    got = O.deform_rhs_csr(h, free_of, con_of, pos, Ffree, Fc, Xc)
    nv = h["nv"]
    F = np.zeros((nv, 9 * P))
    col = np.full(nv, -1, np.int32)
    for v in range(nv):
        if free_of[v] >= 0:
            F[v] = Ffree[free_of[v]]
        else:
            F[v] = Fc[con_of[v]]
            col[v] = con_of[v]
    dense = O.deform_rhs_dense(h, col, pos, F, Xc)
    want = dense[np.flatnonzero(free_of >= 0)]
    assert np.array_equal(got, want)


def test_two_stage_solve_reproduces_a_rigid_motion():
    """F_c = R and X_c = R x_c + t on every sample: the solution is R x + t
    at every free vertex (Lh rows sum to zero)."""
    h, con, pos = _patch(5, n=9)
    r = O.cut_laplacian_csr(h, con, pos)
    ptr, col, val = r["A_ff"]
    free_of, con_of = r["free_of"], r["con_of"]
    nc = int(con.sum())
    ang = 0.7
    R = np.array([[math.cos(ang), -math.sin(ang), 0], [math.sin(ang), math.cos(ang), 0],
                  [0, 0, 1.0]]) @ np.array([[1, 0, 0], [0, math.cos(0.3), -math.sin(0.3)],
                                            [0, math.sin(0.3), math.cos(0.3)]])
    t = np.array([0.3, -1.2, 2.0])
    cidx = np.flatnonzero(con_of >= 0)
    Fc = np.tile(R.reshape(1, 9), (nc, 1))
    Xc = np.zeros((nc, 3))
    Xc[con_of[cidx]] = pos[cidx] @ R.T + t
    b1 = O.harmonic_rhs_csr(h, free_of, con_of, pos, Fc)
    Fv = np.stack([O.pcg(ptr, col, val, b1[:, j], 1e-14 * np.linalg.norm(b1[:, j]))[0]
                   for j in range(9)], 1)
    b2 = O.deform_rhs_csr(h, free_of, con_of, pos, Fv, Fc, Xc)
    Xv = np.stack([O.pcg(ptr, col, val, b2[:, j], 1e-14 * np.linalg.norm(b2[:, j]))[0]
                   for j in range(3)], 1)
    fidx = np.flatnonzero(free_of >= 0)
    want = pos[fidx] @ R.T + t
    assert np.abs(Xv[free_of[fidx]] - want).max() < 1e-9
