"""The reference's own known-answer tests, run against the FP64 oracle.

Each test names the native_decide example(s) it ports.  These pin the
oracle to the spec before it is trusted as the GPU checker."""
import math

import numpy as np
import pytest

import oracle as O
from kat_data import (ATRI, AWIDE, COLOR_GRAPHS, DIAG234, EQ_TRI, FULLPAT, IDENTITY3,
                      LAPLACIAN4, LAPLACIAN4_DENSE, PATH3, PENTAGON, QUAD, SINGLE_TRI, SPD2,
                      STAR4, STRIP, STRIP_POS, TET, TRI3, TRI_PLUS_LEAF, TRIANGLE,
                      TWO_TRI_STRIP, UNIT_SQUARE, WIDEPAT, adj_to_csr, csr_np, he)


def close(a, b, eps):
    return bool(np.all(np.abs(np.asarray(a) - np.asarray(b)) < eps))


# ---------------------------------------------------------- Halfedge.lean --
@pytest.mark.parametrize("m", [TRIANGLE, QUAD, TWO_TRI_STRIP])
def test_hand_meshes_manifold(m):                       # Halfedge.lean:202-204
    assert O.he_manifold(he(m)) == 15


def test_hand_mesh_counts():                            # Halfedge.lean:208-215
    assert he(TRIANGLE)["nh"] == 6 and he(QUAD)["nh"] == 8
    t, q = he(TRIANGLE), he(QUAD)
    assert t["nv"] + (t["nf"] + 1) - t["nh"] // 2 == 2
    assert q["nv"] + (q["nf"] + 1) - q["nh"] // 2 == 2


# --------------------------------------------------- HalfedgeBuilder.lean --
@pytest.mark.parametrize("spec", [True, False])
@pytest.mark.parametrize("case,nh", [(SINGLE_TRI, 6), (STRIP, 10), (TET, 12)])
def test_from_triangles_counts(case, nh, spec):         # HalfedgeBuilder.lean:115-136
    nv, tris = case
    h = O.he_from_triangles(nv, tris, spec=spec)
    assert h["nv"] == nv and h["nf"] == len(tris) and h["nh"] == nh
    assert O.he_manifold(h) == 15


def test_tet_twin_involutive():                         # HalfedgeBuilder.lean:139
    h = O.he_from_triangles(*TET, spec=True)
    assert (h["twin"] >= 0).all()
    assert (h["twin"][h["twin"]] == np.arange(12)).all()


# ------------------------------------------------------ CutAlgorithm.lean --
def loop_len(h, start):
    cur, steps = h["next"][start], 1
    while cur != start and steps < h["nh"]:
        cur = h["next"][cur]
        steps += 1
    return steps


def euler(h):
    return h["nv"] + (h["nf"] + 1) - h["nh"] // 2


def test_subdivide_edge():                              # CutAlgorithm.lean:270-320
    t = O.subdivide_edge(he(TRIANGLE), 0)
    assert (t["nv"], t["nh"], t["nf"]) == (4, 8, 1) and O.he_manifold(t) == 15
    assert loop_len(t, 0) == 4
    q = O.subdivide_edge(he(QUAD), 0)
    assert (q["nv"], q["nh"]) == (5, 10) and O.he_manifold(q) == 15
    t2 = O.subdivide_edge(O.subdivide_edge(he(TRIANGLE), 0), 1)
    assert (t2["nv"], t2["nh"]) == (5, 10) and O.he_manifold(t2) == 15
    assert euler(he(TRIANGLE)) == euler(t)


def test_split_face():                                  # CutAlgorithm.lean:322-367
    q = O.split_face(he(QUAD), 0, 2)
    assert (q["nv"], q["nh"], q["nf"]) == (4, 10, 2) and O.he_manifold(q) == 15
    assert loop_len(q, 1) == 3 and loop_len(q, 0) == 3
    assert euler(he(QUAD)) == euler(q)
    m = O.split_face(O.subdivide_edge(he(QUAD), 0), 1, 3)
    assert (m["nv"], m["nh"], m["nf"]) == (5, 12, 2) and O.he_manifold(m) == 15


def test_insert_crack():                                # CutAlgorithm.lean:419-450
    t = O.insert_crack(he(TRIANGLE), 0)
    assert (t["nv"], t["nh"], t["nf"]) == (4, 8, 1) and O.he_manifold(t) == 15
    assert t["face"][6] == 0 and t["face"][7] == 0
    assert t["twin"][6] == 7 and t["twin"][7] == 6
    assert loop_len(t, 0) == 5


def test_chain_across_two_faces():                      # CutAlgorithm.lean:480-507
    m1 = O.subdivide_edge(he(TWO_TRI_STRIP), 1)
    m2 = O.split_face(m1, 2, 1)
    m3 = O.split_face(m2, 3, 4)
    assert (m3["nv"], m3["nh"], m3["nf"]) == (5, 16, 4)
    assert O.he_manifold(m3) == 15
    assert loop_len(m3, 0) == 3


# --------------------------------------------------- PolygonLaplacian.lean --
def test_triangle_laplacian():                          # PolygonLaplacian.lean:134-158
    L = O.triangle_cot_laplacian(*EQ_TRI)
    assert np.all(np.abs(L.sum(1)) < 1e-12)
    assert np.all(np.abs(L - L.T) < 1e-12)
    assert abs(L[0, 1] - (-1.0 / (2.0 * math.sqrt(3.0)))) < 1e-12


@pytest.mark.parametrize("poly,eps", [(UNIT_SQUARE, 1e-12), (PENTAGON, 1e-9)])
def test_polygon_laplacian(poly, eps):                  # PolygonLaplacian.lean:160-178
    L = O.polygon_cot_laplacian(poly)
    assert np.all(np.abs(L.sum(1)) < eps) and np.all(np.abs(L - L.T) < eps)


# --------------------------------------------------- CutMeshLaplacian.lean --
@pytest.mark.parametrize("m,pos,n", [(TRIANGLE, EQ_TRI, 3), (QUAD, UNIT_SQUARE, 4),
                                     (TWO_TRI_STRIP, STRIP_POS, 4)])
def test_vtlhv_uncut(m, pos, n):                        # CutMeshLaplacian.lean:113-169
    M = O.assemble_vtlhv_dense(he(m), np.zeros(m["nv"], np.uint8), pos)
    assert M.shape == (n, n)
    eps = 1e-12 if m is not TWO_TRI_STRIP else 1e-10
    assert np.all(np.abs(M.sum(1)) < eps) and np.all(np.abs(M - M.T) < eps)


def test_vtlhv_triangle_with_sample():                  # CutMeshLaplacian.lean:172-179
    M = O.assemble_vtlhv_dense(he(TRIANGLE), np.array([1, 0, 0], np.uint8), EQ_TRI)
    assert M.shape == (3, 3) and np.all(np.abs(M - M.T) < 1e-10)
    assert np.all(M[0] == 0) and np.all(M[:, 0] == 0)   # promoted row (HarmonicSolve.lean:14-16)


# ------------------------------------------------------ HarmonicSolve.lean --
@pytest.mark.parametrize("Fc", [[5.0, 10.0, 15.0], [1.0, 0, 0, 0, 1.0, 0, 0, 0, 1.0], [7.0],
                                [1.0, 2.0, 3.0]])
def test_harmonic_solve_triangle_with_sample(Fc):       # HarmonicSolve.lean:92-154
    h = he(TRIANGLE)
    lhs = O.assemble_vtlhv_dense(h, np.array([1, 0, 0], np.uint8), EQ_TRI)
    Fc = np.asarray(Fc).reshape(1, -1)
    rhs = O.harmonic_rhs_dense(h, np.array([0, -1, -1], np.int32), EQ_TRI, Fc)
    X = O.dense_solve_multi(lhs, rhs)
    assert close(X[1], Fc[0], 1e-9) and close(X[2], Fc[0], 1e-9)
    assert np.all(X[0] == 0)                            # caller overlays the sample


# ------------------------------------------------------- SparseLinAlg.lean --
def test_spmv_identity():                               # SparseLinAlg.lean:183-186
    assert close(O.spmv(*csr_np(IDENTITY3), [1.0, 2.0, 3.0]), [1.0, 2.0, 3.0], 1e-12)


@pytest.mark.parametrize("A,b,expect,maxit,eps", [
    (IDENTITY3, [1.0, 2.0, 3.0], [1.0, 2.0, 3.0], 10, 1e-9),         # :189-192
    (DIAG234, [4.0, 9.0, 16.0], [2.0, 3.0, 4.0], 10, 1e-9),          # :203-207
    (SPD2, [3.0, 3.0], [1.0, 1.0], 50, 1e-8),                        # :218-221
    (LAPLACIAN4, [1.0, 0.0, 0.0, 0.0], [0.8, 0.6, 0.4, 0.2], 200, 1e-7),  # :243-246
])
def test_cg_kats(A, b, expect, maxit, eps):
    x, _ = O.cg(*csr_np(A), b, maxit, 1e-12)
    assert close(x, expect, eps)


def test_cg_matches_dense():                            # SparseLinAlg.lean:251-261
    b = [1.0, 2.0, 3.0, 4.0]
    x, _ = O.cg(*csr_np(LAPLACIAN4), b, 200, 1e-12)
    assert close(x, O.dense_solve(LAPLACIAN4_DENSE, b), 1e-7)


def test_cg_with_guess():                               # SparseLinAlg.lean:266-287
    b = [1.0, 2.0, 3.0, 4.0]
    xc, _ = O.cg(*csr_np(LAPLACIAN4), b, 200, 1e-12)
    xw, _ = O.cg(*csr_np(LAPLACIAN4), b, 200, 1e-12, x0=np.zeros(4))
    assert close(xc, xw, 1e-9)
    xt = np.array([0.8, 0.6, 0.4, 0.2])
    x0, it = O.cg(*csr_np(LAPLACIAN4), [1.0, 0, 0, 0], 0, 1e-9, x0=xt)
    assert it == 0 and close(x0, xt, 1e-9)
    xg, _ = O.cg(*csr_np(LAPLACIAN4), b, 200, 1e-12, x0=np.array([3.9, 6.9, 8.9, 6.9]))
    assert close(xc, xg, 1e-7)


def test_solver_kernels_feature():                      # Features/KnownKnowns/SolverKernels.lean:9-18
    ptr, col, val = csr_np(TRI3)
    b = np.array([1.0, 0.0, -1.0])
    x, _ = O.cg(ptr, col, val, b, 100, 1e-12)
    r = b - O.spmv(ptr, col, val, x)
    assert O.dot(r, r) < 1e-9


# ------------------------------------------------ IncompleteCholesky.lean --
def test_ic0_atri():                                    # IncompleteCholesky.lean:144-171
    L = O.ic0_factor3(ATRI, FULLPAT)
    assert abs(L[0, 1]) < 1e-12 and abs(L[0, 2]) < 1e-12 and abs(L[1, 2]) < 1e-12
    assert close(L @ L.T, ATRI, 1e-12)
    b = np.array([3.0, 2.0, 5.0])
    y = np.linalg.solve(L, b)
    x = np.linalg.solve(L.T, y)
    assert close(ATRI @ x, b, 1e-10)


def test_ic0_awide():                                   # IncompleteCholesky.lean:183-186
    L = O.ic0_factor3(AWIDE, WIDEPAT)
    assert close(L @ L.T, AWIDE, 1e-9)


def test_ic0_csr_equals_factor3():
    """The CSR form reproduces factor3 bit for bit on the spec's examples
    (pattern = structural nonzeros of A)."""
    from kat_data import dense_to_csr
    for A, pat in ((ATRI, FULLPAT), (AWIDE, WIDEPAT)):
        L3 = O.ic0_factor3(A, pat)
        lp, lc, lv = O.ic0_factor(*dense_to_csr(A))
        Lc = np.zeros((3, 3))
        for i in range(3):
            for p in range(lp[i], lp[i + 1]):
                Lc[i, lc[p]] = lv[p]
        assert np.array_equal(Lc, L3)


# --------------------------------------------------------- DenseLinAlg.lean --
@pytest.mark.parametrize("A,b,x", [
    (np.eye(3), [1.0, 2.0, 3.0], [1.0, 2.0, 3.0]),
    (np.diag([2.0, 3.0, 4.0]), [4.0, 9.0, 16.0], [2.0, 3.0, 4.0]),
    (np.array([[2.0, 1.0], [1.0, 2.0]]), [3.0, 3.0], [1.0, 1.0]),
    (np.array([[0.0, 1.0], [1.0, 0.0]]), [2.0, 3.0], [3.0, 2.0]),
])
def test_dense_solve_and_lu(A, b, x):                   # DenseLinAlg.lean:298-392
    assert close(O.dense_solve(A, b), x, 1e-12)
    a, piv = O.lu_factor(A)
    assert close(O.lu_solve(a, piv, b), x, 1e-12)


def test_dense_residual_examples():                     # DenseLinAlg.lean:330-339, 395-409
    A = np.array([[2.0, 0, 1], [0, 3, 0], [1, 0, 4]])
    b = np.array([5.0, 6.0, 9.0])
    assert close(A @ O.dense_solve(A, b), b, 1e-10)
    A2 = np.array([[2.0, 1, 0], [1, 2, 1], [0, 1, 2]])
    a, piv = O.lu_factor(A2)
    assert close(A2 @ O.lu_solve(a, piv, b), b, 1e-10)


def test_cholesky():                                    # DenseLinAlg.lean:440-489
    L = O.cholesky_factor(np.array([[4.0, 2.0], [2.0, 3.0]]))
    assert close(L, [[2.0, 0.0], [1.0, math.sqrt(2.0)]], 1e-12)
    assert close(O.cholesky_solve(L, [6.0, 5.0]), [1.0, 1.0], 1e-12)


# ---------------------------------------- HierarchicalSparsifyCompensate.lean --
def test_hsc_path3():                                   # :143-151
    g = O.hsc_eliminate_vertex(PATH3, 1)
    assert len(g) == 1 and abs(O.hsc_edge_weight(g, 0, 2) - 2.0 / 3.0) < 1e-12


def test_hsc_star4():                                   # :164-181
    g = O.hsc_eliminate_vertex(STAR4, 0)
    assert len(g) == 3
    assert abs(O.hsc_edge_weight(g, 1, 2) - 1.0 / 3.0) < 1e-12
    assert abs(O.hsc_edge_weight(g, 1, 3) - 0.5) < 1e-12
    assert abs(O.hsc_edge_weight(g, 2, 3) - 1.0) < 1e-12
    assert abs(O.hsc_degree_weight(g, 1) - 5.0 / 6.0) < 1e-12


def test_hsc_tri_plus_leaf():                           # :194-200
    g = O.hsc_eliminate_vertex(TRI_PLUS_LEAF, 0)
    assert abs(O.hsc_edge_weight(g, 1, 2) - (1.0 + 1.0 / 6.0)) < 1e-12
    assert abs(O.hsc_edge_weight(g, 1, 3) - 4.0 / 6.0) < 1e-12


def test_hsc_isolated_and_sequential():                 # :205-221
    g = O.hsc_eliminate_vertex([(0, 1, 1.5)], 2)
    assert len(g) == 1 and abs(O.hsc_edge_weight(g, 0, 1) - 1.5) < 1e-12
    g2 = O.hsc_eliminate_vertex(O.hsc_eliminate_vertex(PATH3, 1), 0)
    assert len(g2) == 0


# ------------------------------------------------------ GraphColoring.lean --
@pytest.mark.parametrize("name", sorted(COLOR_GRAPHS))
def test_greedy_color(name):                            # GraphColoring.lean:84-113
    adj, ncol = COLOR_GRAPHS[name]
    ptr, col = adj_to_csr(adj)
    color, nc = O.greedy_color(ptr, col)
    assert nc == ncol
    for i, row in enumerate(adj):
        for j in row:
            assert color[i] != color[j]


# ------------------------------------------------- SurfaceProjection.lean --
_ORIGIN, _XU, _YU = (0.0, 0.0, 0.0), (1.0, 0.0, 0.0), (0.0, 1.0, 0.0)
_TRI_MESH = np.array([_ORIGIN, _XU, _YU])


def test_projection_knot_at_vertex():                   # SurfaceProjection.lean:99-101
    r = O.project_to_vertices([_ORIGIN], _TRI_MESH)
    assert r.shape == (1,) and r[0] == 0


@pytest.mark.parametrize("knot,want", [((0.9, 0.0, 0.0), 1), ((0.0, 0.9, 0.0), 2)])
def test_projection_nearest(knot, want):                # SurfaceProjection.lean:105-112
    assert O.project_to_vertices([knot], _TRI_MESH)[0] == want


def test_projection_all_vertex_kind():                  # SurfaceProjection.lean:115-120
    # projectToVertices produces only the vertex kind; every knot gets a
    # vertex (meshIndex >= 0), so every projection is of that kind
    r = O.project_to_vertices([_ORIGIN, _XU, _YU], _TRI_MESH)
    assert r.shape == (3,) and (r >= 0).all() and list(r) == [0, 1, 2]


def test_projection_empty_knots():                      # SurfaceProjection.lean:123-125
    assert O.project_to_vertices(np.zeros((0, 3)), _TRI_MESH).shape == (0,)


def test_projection_empty_mesh():                       # SurfaceProjection.lean:128-130
    r = O.project_to_vertices([_ORIGIN], np.zeros((0, 3)))
    assert r[0] == -1
    # the position falls back to the knot itself (SurfaceProjection.lean:58)
    assert np.array_equal(O.projected_positions([(0.3, 0.2, 0.1)], np.zeros((0, 3)), r),
                          [[0.3, 0.2, 0.1]])
    assert np.array_equal(O.projected_positions([(0.9, 0.0, 0.0)], _TRI_MESH, [1]), [_XU])


@pytest.mark.parametrize("knots,cols", [
    ([_ORIGIN, (0.01, 0.0, 0.0)], [0, 0]),              # SurfaceProjection.lean:133-137
    ([_XU, _YU, _ORIGIN], [0, 1, 2]),                   # SurfaceProjection.lean:141-144
    ([_YU, _ORIGIN], [0, 1]),                           # SurfaceProjection.lean:148-151
    ([_ORIGIN, _XU, _ORIGIN], [0, 1, 0]),               # SurfaceProjection.lean:154-157
])
def test_promote_vertex_samples(knots, cols):
    kc, vc, nc = O.promote_vertex_samples(3, O.project_to_vertices(knots, _TRI_MESH))
    assert list(kc) == cols and nc == max(cols) + 1
