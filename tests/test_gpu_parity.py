"""GPU parity: the gfx950 library against the FP64 oracle on identical inputs.

Integer/index outputs and the FP64 assembly, SpMV, IC(0) factor and apply
are compared bit for bit (np.array_equal); FP32 kernels and solves against
the stated tolerances."""
import numpy as np
import pytest

import oracle as O
from kat_data import (DIAG234, EQ_TRI, IDENTITY3, LAPLACIAN4, QUAD, SINGLE_TRI, SPD2, STRIP,
                      TET, TRI3, TRIANGLE, csr_np, he)

pytestmark = pytest.mark.gpu


def _mesh(name):
    from tool_godot_curvenet_amd import meshgen as G
    return G.config_mesh(name)


def _random_mesh(seed, n=6):
    """Jittered n x n grid patch (open boundary): small enough for the
    literal O(F^2) spec builder."""
    rng = np.random.default_rng(seed)
    tris = []
    for r in range(n - 1):
        for c in range(n - 1):
            a, b, d, e = r * n + c, r * n + c + 1, (r + 1) * n + c, (r + 1) * n + c + 1
            if rng.random() < 0.5:
                tris += [(a, b, e), (a, e, d)]
            else:
                tris += [(a, b, d), (b, e, d)]
    order = rng.permutation(len(tris))
    tris = np.asarray(tris, np.int32)[order]
    y, x = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    pos = np.stack([x.ravel(), y.ravel(), np.zeros(n * n)], 1) + rng.uniform(-0.2, 0.2, (n * n, 3))
    return n * n, tris, pos


# ------------------------------------------------------------- halfedges --
@pytest.mark.parametrize("case", [SINGLE_TRI, STRIP, TET])
def test_halfedges_spec_kats(gpu, case):
    from tool_godot_curvenet_amd import from_triangles
    nv, tris = case
    g = from_triangles(nv, tris)
    o = O.he_from_triangles(nv, tris, spec=True)
    for k in ("target", "twin", "next", "face"):
        assert np.array_equal(g[k], o[k]), k


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_halfedges_random_patch(gpu, seed):
    from tool_godot_curvenet_amd import from_triangles
    nv, tris, _ = _random_mesh(seed)
    g = from_triangles(nv, tris)
    o = O.he_from_triangles(nv, tris, spec=True)
    for k in ("target", "twin", "next", "face"):
        assert np.array_equal(g[k], o[k]), k
    assert O.he_manifold(g) == 15


@pytest.mark.parametrize("name", ["cylinder64", "capsules100k"])
def test_halfedges_config_meshes(gpu, name):
    from tool_godot_curvenet_amd import from_triangles
    m = _mesh(name)
    g = from_triangles(m.nv, m.tris)
    o = O.he_from_triangles(m.nv, m.tris)
    for k in ("target", "twin", "next", "face"):
        assert np.array_equal(g[k], o[k]), k


# -------------------------------------------------------------- seam cut --
@pytest.mark.parametrize("name", ["cylinder64", "icosphere6", "capsules100k"])
def test_cut_seams_bit_exact(gpu, name):
    from tool_godot_curvenet_amd import cut_seams
    m = _mesh(name)
    nt, origin, con, nv2 = cut_seams(m.nv, m.tris, m.cut_edges)
    ont, oorigin, ocon, onv2 = O.cut_edges(m.nv, m.tris, m.cut_edges)
    assert nv2 == onv2
    assert np.array_equal(nt, ont)
    assert np.array_equal(origin, oorigin)
    assert np.array_equal(con, ocon)


def test_cut_cylinder_closed_form(gpu):
    """Seam ring 32 of the 64x64 cylinder: 64 copies at nv + j in ring order."""
    from tool_godot_curvenet_amd import cut_seams
    m = _mesh("cylinder64")
    nt, origin, con, nv2 = cut_seams(m.nv, m.tris, m.cut_edges)
    assert nv2 == 4096 + 64
    assert np.array_equal(origin[4096:], 32 * 64 + np.arange(64))
    assert con.sum() == 128


def test_cut_bad_edge_rejected(gpu):
    from tool_godot_curvenet_amd import cut_seams
    from tool_godot_curvenet_amd._lib import CurvenetError
    m = _mesh("cylinder64")
    with pytest.raises(CurvenetError):
        cut_seams(m.nv, m.tris, np.array([[0, 2000]], np.int32))


# -------------------------------------------------------------- Laplacian --
def _system(name):
    from tool_godot_curvenet_amd import CutLaplacian
    m = _mesh(name)
    lap = CutLaplacian.from_mesh(m)
    return m, lap


def _oracle_system(lap):
    h = O.he_from_triangles(lap.pos.shape[0], lap.tris)
    return h, O.cut_laplacian_csr(h, lap.constrained, lap.pos)


@pytest.mark.parametrize("name", ["cylinder64", "icosphere6"])
def test_cut_laplacian_bit_exact(gpu, name):
    m, lap = _system(name)
    h, o = _oracle_system(lap)
    assert lap.n_free == o["n_free"]
    A = lap.A_ff()
    for got, want in zip((A.row_ptr, A.col_idx, A.values), o["A_ff"]):
        assert np.array_equal(got, want)
    B = lap.A_fc()
    for got, want in zip((B.row_ptr, B.col_idx, B.values), o["A_fc"]):
        assert np.array_equal(got, want)
    free_of, con_of = lap.maps()
    assert np.array_equal(free_of, o["free_of"]) and np.array_equal(con_of, o["con_of"])
    he_g = lap.halfedges()
    for k in ("target", "twin", "next", "face"):
        assert np.array_equal(he_g[k], h[k])


@pytest.mark.parametrize("name", ["cylinder64", "icosphere6"])
def test_reassembly_bit_exact(gpu, name):
    """New positions through set_positions (the assembly plan built with the
    system, summed again): A_ff, A_fc and the FP32 mirror equal the oracle's
    assembly at those positions and a freshly built system's."""
    from tool_godot_curvenet_amd import CutLaplacian
    from tool_godot_curvenet_amd.curvenet import _np
    m, lap = _system(name)
    rng = np.random.default_rng(21)
    pos2 = lap.pos + rng.uniform(-0.01, 0.01, size=lap.pos.shape)
    lap.set_positions(pos2)
    h = O.he_from_triangles(pos2.shape[0], lap.tris)
    o = O.cut_laplacian_csr(h, lap.constrained, pos2)
    A, B = lap.A_ff(), lap.A_fc()
    for got, want in zip((A.row_ptr, A.col_idx, A.values), o["A_ff"]):
        assert np.array_equal(got, want)
    for got, want in zip((B.row_ptr, B.col_idx, B.values), o["A_fc"]):
        assert np.array_equal(got, want)
    v32 = _np(lap.info.d_ff_val32, lap.info.nnz_ff, np.float32)
    assert np.array_equal(v32, A.values.astype(np.float32))
    fresh = CutLaplacian(pos2, lap.tris, lap.constrained)
    assert np.array_equal(fresh.A_ff().values, A.values)
    assert np.array_equal(fresh.A_fc().values, B.values)


def test_cut_laplacian_matches_dense_spec(gpu):
    """Small cut patch: GPU CSR == the literal dense assembleVtLhV."""
    from tool_godot_curvenet_amd import CutLaplacian, cut_seams
    nv, tris, pos = _random_mesh(7, n=7)
    # cut along the middle row, interior edges only
    cut = np.array([(3 * 7 + c, 3 * 7 + c + 1) for c in range(1, 5)], np.int32)
    nt, origin, con, nv2 = cut_seams(nv, tris, cut)
    lap = CutLaplacian(pos[origin], nt, con)
    h = O.he_from_triangles(nv2, nt, spec=True)
    D = O.assemble_vtlhv_dense(h, con, pos[origin])
    free = np.flatnonzero(con == 0)
    Aff = D[np.ix_(free, free)]
    A = lap.A_ff().dense()
    assert np.array_equal(A, Aff)


@pytest.mark.parametrize("name", ["cylinder64", "icosphere6"])
def test_harmonic_rhs_bit_exact(gpu, name):
    from tool_godot_curvenet_amd import meshgen as G
    m, lap = _system(name)
    h, o = _oracle_system(lap)
    Xc = G.seeded_boundary(lap.n_con, 3)
    got = lap.harmonic_rhs(Xc)
    want = O.harmonic_rhs_csr(h, o["free_of"], o["con_of"], lap.pos, Xc)
    assert np.array_equal(got, want)


# ------------------------------------------------------------ SpMV / SpMM --
@pytest.mark.parametrize("k", [1, 3, 8])
def test_spmm_f64_bit_exact(gpu, k):
    from tool_godot_curvenet_amd import spmm
    m, lap = _system("cylinder64")
    A = lap.A_ff()
    X = np.random.default_rng(k).standard_normal((A.rows, k))
    got = spmm(A, X, "f64")
    want = O.spmm(A.row_ptr, A.col_idx, A.values, X)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("k", [1, 3, 4, 64, 256])
def test_spmm_f32(gpu, k):
    from tool_godot_curvenet_amd import spmm
    m, lap = _system("cylinder64")
    A = lap.A_ff()
    X = np.random.default_rng(k).standard_normal((A.rows, k)).astype(np.float32)
    got = spmm(A, X, "f32").astype(np.float64)
    want = O.spmm(A.row_ptr, A.col_idx, A.values, X.astype(np.float64))
    absA = O.spmm(A.row_ptr, A.col_idx, np.abs(A.values), np.abs(X).astype(np.float64))
    # FP32 values and FMA accumulation: |err| <= c * eps32 * (|A||x|)
    assert np.all(np.abs(got - want) <= 1e-5 * absA + 1e-30)


def test_spmv_kat(gpu):
    from tool_godot_curvenet_amd import CsrMatrix, spmv
    p, c, v = csr_np(IDENTITY3)
    A = CsrMatrix(3, 3, p, c, v)
    assert np.array_equal(spmv(A, [1.0, 2.0, 3.0]), [1.0, 2.0, 3.0])


# ------------------------------------------------------------------ IC(0) --
@pytest.mark.parametrize("ordering", ["color", "natural", "block"])
def test_ic0_factor_bit_exact(gpu, ordering):
    from tool_godot_curvenet_amd import IC0
    m, lap = _system("cylinder64")
    f = IC0(lap, ordering)
    A = lap.A_ff()
    perm = f.perm()
    if ordering == "color":
        color, nc = O.greedy_color(A.row_ptr, A.col_idx)
        assert f.info.nlevels == nc
        assert np.array_equal(perm, np.argsort(color, kind="stable"))
    elif ordering == "natural":
        assert np.array_equal(perm, np.arange(A.rows)) or f.info.nlevels > 1
    else:
        assert f.info.max_block <= 32 and f.info.nblocks >= A.rows // 32
        assert np.array_equal(np.sort(perm), np.arange(A.rows))
    pa = f.permuted_matrix()
    optr, ocol, oval = O.permute_csr(A.row_ptr, A.col_idx, A.values, perm)
    assert np.array_equal(pa.row_ptr, optr) and np.array_equal(pa.col_idx, ocol)
    assert np.array_equal(pa.values, oval)
    assert f.info.shift == 0.0
    lp, lc, lv = O.ic0_factor(optr, ocol, oval)
    gp, gc, gv, gd = f.factor()
    # oracle keeps the diagonal last in each row; the GPU stores it apart
    n = A.rows
    strict = np.concatenate([np.arange(lp[i], lp[i + 1] - 1) for i in range(n)])
    assert np.array_equal(gc, lc[strict]) and np.array_equal(gv, lv[strict])
    assert np.array_equal(gd, lv[lp[1:] - 1])


@pytest.mark.parametrize("ordering", ["color", "block"])
def test_ic0_apply_bit_exact(gpu, ordering):
    from tool_godot_curvenet_amd import IC0
    m, lap = _system("cylinder64")
    f = IC0(lap, ordering)
    A = lap.A_ff()
    perm = f.perm()
    optr, ocol, oval = O.permute_csr(A.row_ptr, A.col_idx, A.values, perm)
    lp, lc, lv = O.ic0_factor(optr, ocol, oval)
    r = np.random.default_rng(5).standard_normal((A.rows, 3))
    got = f.apply(r)
    for c in range(3):
        want_p = O.ic0_solve(lp, lc, lv, r[perm, c])
        want = np.empty_like(want_p)
        want[perm] = want_p
        assert np.array_equal(got[:, c], want)


# -------------------------------------------------------------------- PCG --
def _relres(A, X, B):
    R = B - O.spmm(A.row_ptr, A.col_idx, A.values, X)
    return np.linalg.norm(R, axis=0) / np.linalg.norm(B, axis=0)


@pytest.mark.parametrize("precision", ["mixed", "fp64", "fp64_plain"])
@pytest.mark.parametrize("precond", ["ic0-block", "ic0-color", "jacobi"])
@pytest.mark.parametrize("k", [3, 96])
def test_pcg_cylinder(gpu, precision, precond, k):
    from tool_godot_curvenet_amd import IC0
    from tool_godot_curvenet_amd import meshgen as G
    m, lap = _system("cylinder64")
    A = lap.A_ff()
    B = lap.harmonic_rhs(G.seeded_boundary(lap.n_con, k, seed=k))
    ic0 = IC0(lap, precond.split("-")[1]) if precond.startswith("ic0") else None
    x, info = lap.solve(B, precond=precond.split("-")[0], precision=precision, tol=1e-9,
                        ic0=ic0)
    X = x.cpu().numpy()
    assert info.status == 0, info.status
    rr = _relres(A, X, B)
    assert np.all(rr <= 1e-9), rr
    assert info.relres_max <= 1e-9


@pytest.mark.parametrize("name", ["cylinder64", "icosphere6"])
@pytest.mark.parametrize("ordering", ["block", "color", "natural"])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-11), ("fp64_plain", 1e-11),
                                           ("mixed", 2e-5)])
@pytest.mark.parametrize("k", [3, 100])
def test_pcg_first_step_is_ic0_direction(gpu, name, ordering, precision, tol, k):
    """From x0 = 0 one PCG step gives x1 = alpha0 M^-1 b, so x1 / |x1| must
    be the oracle's IC(0) solve of b (on the permuted system, mapped back)
    normalised: this pins the solver's triangular sweeps (every kernel form:
    point and block levels, narrow and wide column tiles) to the oracle."""
    from tool_godot_curvenet_amd import IC0
    from tool_godot_curvenet_amd import meshgen as G
    m, lap = _system(name)
    B = lap.harmonic_rhs(G.seeded_boundary(lap.n_con, k, seed=40 + k))
    ic0 = IC0(lap, ordering)
    x, info = lap.solve(B, precond="ic0", precision=precision, tol=1e-9, ic0=ic0,
                        max_iter=1)
    assert info.iters == 1
    X = x.cpu().numpy()
    perm = ic0.perm()
    Ap = ic0.permuted_matrix()
    lp, lc, lv = O.ic0_factor(Ap.row_ptr, Ap.col_idx, Ap.values)
    worst = 0.0
    for c in range(k):
        z = np.empty(lap.n_free)
        z[perm] = O.ic0_solve(lp, lc, lv, B[perm, c])
        e = np.linalg.norm(X[:, c] / np.linalg.norm(X[:, c]) - z / np.linalg.norm(z))
        worst = max(worst, e)
    assert worst <= tol, worst


@pytest.mark.parametrize("ordering", ["block", "color"])
@pytest.mark.parametrize("k", [7, 100, 300])
def test_pcg_mixed_ragged_k(gpu, ordering, k):
    """Column counts that leave the last column tile partly empty."""
    from tool_godot_curvenet_amd import IC0
    from tool_godot_curvenet_amd import meshgen as G
    m, lap = _system("icosphere6")
    B = lap.harmonic_rhs(G.seeded_boundary(lap.n_con, k, seed=50 + k))
    x, info = lap.solve(B, precond="ic0", precision="mixed", tol=1e-9,
                        ic0=IC0(lap, ordering))
    assert info.status == 0
    assert np.all(_relres(lap.A_ff(), x.cpu().numpy(), B) <= 1e-9)


def test_pcg_solution_matches_oracle(gpu):
    """Positions within 1e-9 relative of the oracle's FP64 IC(0)-PCG."""
    from tool_godot_curvenet_amd import meshgen as G
    m, lap = _system("cylinder64")
    A = lap.A_ff()
    B = lap.harmonic_rhs(G.seeded_boundary(lap.n_con, 3))
    x, info = lap.solve(B, precond="ic0", precision="mixed", tol=1e-12)
    X = x.cpu().numpy()
    lp, lc, lv = O.ic0_factor(A.row_ptr, A.col_idx, A.values)
    for c in range(3):
        xo, it = O.pcg(A.row_ptr, A.col_idx, A.values, B[:, c],
                       1e-13 * np.linalg.norm(B[:, c]), factor=(lp, lc, lv))
        err = np.linalg.norm(X[:, c] - xo) / np.linalg.norm(xo)
        assert err <= 1e-9, err


def test_pcg_wide_k(gpu):
    from tool_godot_curvenet_amd import meshgen as G
    m, lap = _system("cylinder64")
    A = lap.A_ff()
    B = lap.harmonic_rhs(G.seeded_boundary(lap.n_con, 96, seed=3))
    x, info = lap.solve(B, precond="ic0", precision="mixed", tol=1e-9)
    assert info.status == 0
    assert np.all(_relres(A, x.cpu().numpy(), B) <= 1e-9)


def test_pcg_warm_start_zero_iters(gpu):
    """cgWithGuess seeded with the solution converges in 0 iterations
    (SparseLinAlg.lean:272-277)."""
    from tool_godot_curvenet_amd import meshgen as G
    m, lap = _system("cylinder64")
    B = lap.harmonic_rhs(G.seeded_boundary(lap.n_con, 3))
    x, _ = lap.solve(B, tol=1e-11)
    x2, info = lap.solve(B, tol=1e-9, x0=x)
    assert info.iters == 0 and info.status == 0


@pytest.mark.parametrize("A,b,expect,maxit,eps", [
    (IDENTITY3, [1.0, 2.0, 3.0], [1.0, 2.0, 3.0], 10, 1e-9),
    (DIAG234, [4.0, 9.0, 16.0], [2.0, 3.0, 4.0], 10, 1e-9),
    (SPD2, [3.0, 3.0], [1.0, 1.0], 50, 1e-8),
    (LAPLACIAN4, [1.0, 0.0, 0.0, 0.0], [0.8, 0.6, 0.4, 0.2], 200, 1e-7),
])
def test_cg_spec_kats_on_gpu(gpu, A, b, expect, maxit, eps):   # SparseLinAlg.lean:171-246
    from tool_godot_curvenet_amd import CsrMatrix, cg
    p, c, v = csr_np(A)
    x = cg(CsrMatrix(len(p) - 1, len(p) - 1, p, c, v), b, maxit, 1e-12)
    assert np.all(np.abs(x - np.asarray(expect)) < eps)


def test_cg_solver_kernels_feature_on_gpu(gpu):                # SolverKernels.lean:9-18
    from tool_godot_curvenet_amd import CsrMatrix, cg
    p, c, v = csr_np(TRI3)
    b = np.array([1.0, 0.0, -1.0])
    x = cg(CsrMatrix(3, 3, p, c, v), b, 100, 1e-12)
    r = b - O.spmv(p, c, v, x)
    assert O.dot(r, r) < 1e-9


def test_harmonic_solve_triangle_with_sample(gpu):             # HarmonicSolve.lean:104-115
    from tool_godot_curvenet_amd import CutLaplacian, harmonic_solve
    lap = CutLaplacian(EQ_TRI, [(0, 1, 2)], [1, 0, 0])
    X, info = harmonic_solve(lap, [[5.0, 10.0, 15.0]], precond="jacobi", precision="fp64")
    assert np.all(np.abs(X[1] - [5.0, 10.0, 15.0]) < 1e-9)
    assert np.all(np.abs(X[2] - [5.0, 10.0, 15.0]) < 1e-9)
    assert np.all(X[0] == 0.0)


# ------------------------------------------------------------- projection --
def test_project_to_vertices(gpu):
    from tool_godot_curvenet_amd import project_to_vertices
    m = _mesh("cylinder64")
    q = np.random.default_rng(3).uniform(-1, 1, (200, 3))
    q[:5] = m.pos[[0, 5, 5, 77, 4000]]                 # exact hits incl. repeats
    assert np.array_equal(project_to_vertices(q, m.pos), O.project_to_vertices(q, m.pos))


_TRI_MESH = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])


@pytest.mark.parametrize("knots,want,cols", [
    ([[0.0, 0.0, 0.0]], [0], [0]),                                   # SurfaceProjection.lean:99-101
    ([[0.9, 0.0, 0.0]], [1], [0]),                                   # :105-107
    ([[0.0, 0.9, 0.0]], [2], [0]),                                   # :110-112
    ([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]], [0, 1, 2], [0, 1, 2]),  # :115-120, 141-144
    ([[0.0, 0.0, 0.0], [0.01, 0.0, 0.0]], [0, 0], [0, 0]),           # :133-137
    ([[0.0, 1.0, 0.0], [0.0, 0.0, 0.0]], [2, 0], [0, 1]),            # :148-151
    ([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 0.0]], [0, 1, 0], [0, 1, 0]),  # :154-157
])
def test_projection_spec_kats(gpu, knots, want, cols):
    from tool_godot_curvenet_amd import project_to_vertices, promote_vertex_samples
    idx = project_to_vertices(knots, _TRI_MESH)
    assert list(idx) == want
    kc, vc, nc = promote_vertex_samples(3, idx)
    assert list(kc) == cols and nc == max(cols) + 1
    okc, ovc, onc = O.promote_vertex_samples(3, idx)
    assert np.array_equal(kc, okc) and np.array_equal(vc, ovc) and nc == onc


def test_projection_empty(gpu):                              # SurfaceProjection.lean:123-130
    from tool_godot_curvenet_amd import project_to_vertices, promote_vertex_samples
    assert project_to_vertices(np.zeros((0, 3)), _TRI_MESH).shape == (0,)
    idx = project_to_vertices([[0.0, 0.0, 0.0]], np.zeros((0, 3)))
    assert list(idx) == [-1]
    kc, vc, nc = promote_vertex_samples(0, idx)
    assert list(kc) == [-1] and vc.shape == (0,) and nc == 0
    kc, vc, nc = promote_vertex_samples(4, np.zeros(0, np.int32))
    assert kc.shape == (0,) and list(vc) == [-1] * 4 and nc == 0


def test_projection_ties(gpu):
    """Duplicated vertices and equidistant knots: the lowest index wins."""
    from tool_godot_curvenet_amd import project_to_vertices
    rng = np.random.default_rng(11)
    pos = rng.uniform(-1, 1, (300, 3))
    pos[200:250] = pos[10:60]
    g = np.array([[0.0, 0.0, 0.0], [0.5, 0.0, 0.0], [0.0, 0.5, 0.0], [0.5, 0.5, 0.0]]) + 3.0
    pos = np.concatenate([pos, g])
    q = np.concatenate([rng.uniform(-1, 1, (100, 3)), pos[10:60],
                        [[3.25, 3.0, 3.0], [3.0, 3.25, 3.0], [3.25, 3.25, 3.0]]])
    got = project_to_vertices(q, pos)
    assert np.array_equal(got, O.project_to_vertices(q, pos))
    assert (got[100:150] == np.arange(10, 60)).all() and list(got[-3:]) == [300] * 3


def test_projection_promotion_c2_knots(gpu):
    """The 8 x 32 loop knots of the level-6 icosphere (the C2 curves):
    projection and column assignment bit-exact against the oracle."""
    from tool_godot_curvenet_amd import meshgen, project_to_vertices, promote_vertex_samples
    m = meshgen.geodesic_sphere(64, jitter_frac=0.05, seed=meshgen.BASE_SEED + 2)
    knots = np.concatenate(meshgen.loop_knots(m, 8, 30.0, 32, seed=meshgen.BASE_SEED + 1))
    idx = project_to_vertices(knots, m.pos)
    assert np.array_equal(idx, O.project_to_vertices(knots, m.pos))
    kc, vc, nc = promote_vertex_samples(m.nv, idx)
    okc, ovc, onc = O.promote_vertex_samples(m.nv, idx)
    assert np.array_equal(kc, okc) and np.array_equal(vc, ovc) and nc == onc
    assert nc == len(np.unique(idx))


@pytest.mark.parametrize("seed", range(3))
def test_promote_random(gpu, seed):
    from tool_godot_curvenet_amd import promote_vertex_samples
    rng = np.random.default_rng(seed)
    nv = int(rng.integers(1, 5000))
    idx = rng.integers(-1, nv, int(rng.integers(1, 20000))).astype(np.int32)
    kc, vc, nc = promote_vertex_samples(nv, idx)
    okc, ovc, onc = O.promote_vertex_samples(nv, idx)
    assert np.array_equal(kc, okc) and np.array_equal(vc, ovc) and nc == onc


def test_promote_rejects_out_of_range(gpu):
    from tool_godot_curvenet_amd import promote_vertex_samples
    with pytest.raises(RuntimeError):
        promote_vertex_samples(3, [0, 1, 3])


def _hub_matrix(n, seed):
    """SPD graph Laplacian + 0.01 I: a ring with +-1, +-2 neighbours and 24
    hub vertices joined to 10-20 random others, so some rows hold up to ~25
    entries and most hold 5."""
    from tool_godot_curvenet_amd import CsrMatrix
    rng = np.random.default_rng(seed)
    W = {}
    for i in range(n):
        for d in (1, 2):
            j = (i + d) % n
            W[(min(i, j), max(i, j))] = rng.uniform(0.5, 2.0)
    for h in rng.choice(n, 24, replace=False):
        for j in rng.choice(n, rng.integers(10, 21), replace=False):
            if j != h:
                W[(min(h, j), max(h, j))] = rng.uniform(0.5, 2.0)
    rows = [dict() for _ in range(n)]
    for (i, j), w in W.items():
        rows[i][j] = -w
        rows[j][i] = -w
        rows[i][i] = rows[i].get(i, 0.0) + w
        rows[j][j] = rows[j].get(j, 0.0) + w
    ptr, col, val = [0], [], []
    for i in range(n):
        rows[i][i] += 0.01
        for j in sorted(rows[i]):
            col.append(j)
            val.append(rows[i][j])
        ptr.append(len(col))
    return CsrMatrix(n, n, np.array(ptr, np.int32), np.array(col, np.int32),
                     np.array(val, np.float64))


@pytest.mark.parametrize("precision", ["mixed", "fp64", "fp64_plain"])
@pytest.mark.parametrize("k", [3, 100])
def test_pcg_csr_long_rows(gpu, precision, k):
    """Rows longer than the multi-column SpMM's 8-wide row table take its
    overflow entries; the FP64 residual (plain CSR) checks the result and the
    iteration count stays at the oracle's FP64 Jacobi-PCG count."""
    from tool_godot_curvenet_amd import CsrSystem
    A = _hub_matrix(3000, 7)
    assert np.diff(A.row_ptr).max() > 16
    B = np.random.default_rng(k).normal(size=(A.rows, k))
    x, info = CsrSystem(A).solve(B, precond="jacobi", precision=precision, tol=1e-9)
    assert info.status == 0
    assert np.all(_relres(A, x.cpu().numpy(), B) <= 1e-9)
    _, it = O.pcg(A.row_ptr, A.col_idx, A.values, B[:, 0], 1e-9 * np.linalg.norm(B[:, 0]))
    # FP64 follows the oracle's recurrence; the FP32 inner loop of "mixed"
    # costs some iterations on this poorly conditioned matrix (kappa ~ 1e4)
    lo, hi = (0.9, 1.1) if precision != "mixed" else (0.9, 1.3)
    assert lo * it - 3 <= info.iters <= hi * it + 3, (info.iters, it)


@pytest.mark.parametrize("precision", ["mixed", "fp64", "fp64_plain"])
def test_jacobi_zero_diagonal(gpu, precision):
    """Jacobi.lean:7-10: a zero diagonal entry gives 0, not 1/0.  A 1-D
    Laplacian (+0.1 I) with rows 5 and 17 emptied to an explicit zero
    diagonal and zero right-hand side: Jacobi-PCG keeps those rows at 0
    (z, p and x stay 0 there; 1/0 would make them inf/NaN) and solves the
    rest to the FP64 tolerance."""
    from tool_godot_curvenet_amd import CsrMatrix, CsrSystem
    n, zero = 40, (5, 17)
    ptr, col, val = [0], [], []
    for i in range(n):
        if i in zero:
            col.append(i)
            val.append(0.0)
        else:
            for j, w in ((i - 1, -1.0), (i, 2.1), (i + 1, -1.0)):
                if 0 <= j < n and j not in zero:
                    col.append(j)
                    val.append(w)
        ptr.append(len(col))
    A = CsrMatrix(n, n, np.array(ptr, np.int32), np.array(col, np.int32),
                  np.array(val, np.float64))
    B = np.random.default_rng(4).normal(size=(n, 3))
    B[list(zero)] = 0.0
    x, info = CsrSystem(A).solve(B, precond="jacobi", precision=precision, tol=1e-9)
    X = x.cpu().numpy()
    assert info.status == 0 and np.all(np.isfinite(X))
    assert np.all(X[list(zero)] == 0.0)
    keep = [i for i in range(n) if i not in zero]
    D = A.dense()[np.ix_(keep, keep)]
    ref = np.linalg.solve(D, B[keep])
    assert np.abs(X[keep] - ref).max() <= 1e-8 * np.abs(ref).max()


def test_solve_after_set_positions(gpu):
    """New positions change the matrix values; the next solve uses them."""
    from tool_godot_curvenet_amd import meshgen as G
    m, lap = _system("cylinder64")
    B0 = lap.harmonic_rhs(G.seeded_boundary(lap.n_con, 96, seed=3))
    lap.solve(B0, precond="jacobi", precision="mixed", tol=1e-9)
    rng = np.random.default_rng(11)
    lap.set_positions(lap.pos * rng.uniform(0.8, 1.25, size=(1, 3)))
    A = lap.A_ff()
    B = lap.harmonic_rhs(G.seeded_boundary(lap.n_con, 96, seed=4))
    x, info = lap.solve(B, precond="jacobi", precision="mixed", tol=1e-9)
    assert info.status == 0
    assert np.all(_relres(A, x.cpu().numpy(), B) <= 1e-9)


# The solver's row tables (the LDS window's packed weights, the row table's
# values) are refilled only after an assembly: each form and preconditioner
# solves with the values of the positions set last, before and after a
# second change, and a solve that changes nothing reuses them.
@pytest.mark.parametrize("precond,form", [("jacobi", 0), ("jacobi", 1), ("hsc", 0),
                                          ("hsc", 1)])
def test_solve_tables_follow_set_positions(gpu, precond, form):
    from tool_godot_curvenet_amd import HSC
    from tool_godot_curvenet_amd import meshgen as G
    m, lap = _system("cylinder64")
    lap.set_tune({"spmm_form": form})
    rng = np.random.default_rng(21)
    pos0 = lap.pos.copy()
    for step in range(3):
        if step:
            lap.set_positions(pos0 * rng.uniform(0.8, 1.25, size=(1, 3)))
        A = lap.A_ff()
        B = lap.harmonic_rhs(G.seeded_boundary(lap.n_con, 96, seed=30 + step))
        kw = {"hsc": HSC(lap)} if precond == "hsc" else {}
        for _ in range(2):  # the second solve reuses the tables
            x, info = lap.solve(B, precond=precond, precision="mixed", tol=1e-9, **kw)
            assert info.status == 0
            assert np.all(_relres(A, x.cpu().numpy(), B) <= 1e-9), (step, precond, form)


@pytest.mark.parametrize("k", [1, 2, 3, 4])
def test_spmm_f32_long_rows_and_grid_stride(gpu, k):
    """Thread-per-row SpMM (k <= 4): rows longer than its 8-entry chunk (hub
    rows of up to ~25 entries), empty rows, and a matrix of more rows than
    one grid of threads (the grid-stride walk)."""
    import scipy.sparse as sps
    from tool_godot_curvenet_amd import CsrMatrix, spmm
    A = _hub_matrix(3000, 11)
    p = A.row_ptr.copy()
    # empty rows 100 and 101: drop their entries
    keep = np.ones(len(A.col_idx), bool)
    for r in (100, 101):
        keep[p[r]:p[r + 1]] = False
    rows = np.repeat(np.arange(A.rows), np.diff(p))[keep]
    M = sps.csr_matrix((A.values[keep], (rows, A.col_idx[keep])), shape=(A.rows, A.rows))
    n2 = 2_200_000
    T = sps.diags([np.full(n2 - 1, -1.0), np.full(n2, 2.5), np.full(n2 - 1, -1.0)],
                  [-1, 0, 1], format="csr")
    for Ms in (M, T):
        Ms.sort_indices()
        C = CsrMatrix(Ms.shape[0], Ms.shape[0], Ms.indptr.astype(np.int32),
                      Ms.indices.astype(np.int32), Ms.data)
        X = np.random.default_rng(k).standard_normal((C.rows, k)).astype(np.float32)
        got = spmm(C, X, "f32").astype(np.float64)
        M32 = sps.csr_matrix((Ms.data.astype(np.float32).astype(np.float64), Ms.indices,
                              Ms.indptr), shape=Ms.shape)
        want = M32 @ X.astype(np.float64)
        absA = abs(M32) @ np.abs(X).astype(np.float64)
        assert np.all(np.abs(got - want) <= 1e-5 * absA + 1e-30)
        assert np.all(got[100:102] == 0.0) if Ms is M else True
