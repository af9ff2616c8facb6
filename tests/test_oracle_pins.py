"""Independent checks that pin the oracle beyond the spec's own KATs:
literal-spec vs fast implementations, dense-vs-sparse bit equality,
closed forms, invariants and textbook identities.  Each is chosen so that a
plausible mistake (dropped term, wrong sign, swapped operand, wrong index)
fails it; the planted-mistake tests at the end show two such failures."""
import numpy as np
import pytest

import oracle as O
from tool_godot_curvenet_amd import meshgen as G


def grid_patch(n=7, seed=0, jit=0.2, shuffle=True):
    rng = np.random.default_rng(seed)
    tris = []
    for r in range(n - 1):
        for c in range(n - 1):
            a, b, d, e = r * n + c, r * n + c + 1, (r + 1) * n + c, (r + 1) * n + c + 1
            if rng.random() < 0.5:
                tris += [(a, b, e), (a, e, d)]
            else:
                tris += [(a, b, d), (b, e, d)]
    tris = np.asarray(tris, np.int32)
    if shuffle:
        tris = tris[rng.permutation(len(tris))]
    y, x = np.meshgrid(np.arange(n, dtype=float), np.arange(n, dtype=float), indexing="ij")
    pos = np.stack([x.ravel(), y.ravel(), np.zeros(n * n)], 1)
    pos[:, :2] += rng.uniform(-jit, jit, (n * n, 2))
    return n * n, tris, pos


# ---------------------------------------------------------- halfedge build --
@pytest.mark.parametrize("seed", range(5))
def test_fast_builder_equals_spec_builder(seed):
    nv, tris, _ = grid_patch(8, seed)
    a = O.he_from_triangles(nv, tris, spec=True)
    b = O.he_from_triangles(nv, tris, spec=False)
    for k in ("target", "twin", "next", "face"):
        assert np.array_equal(a[k], b[k])
    assert O.he_manifold(a) == 15


def test_builders_agree_on_nonmanifold_input():
    """A duplicated triangle makes directed edges repeat: the spec takes the
    first match; both builders must pick the same one."""
    tris = [(0, 1, 2), (2, 1, 3), (0, 1, 2), (1, 0, 4)]
    a = O.he_from_triangles(5, tris, spec=True)
    b = O.he_from_triangles(5, tris, spec=False)
    for k in ("target", "twin", "next", "face"):
        assert np.array_equal(a[k], b[k])
    assert O.he_manifold(a) != 15


def test_manifold_detects_planted_defects():
    nv, tris, _ = grid_patch(5, 1)
    h = O.he_from_triangles(nv, tris)
    bad = dict(h, twin=h["twin"].copy())
    bad["twin"][0] = bad["twin"][1]
    assert O.he_manifold(bad) & 2 == 0
    bad = dict(h, next=h["next"].copy())
    bad["next"][0] = bad["next"][3]
    assert O.he_manifold(bad) != 15


# --------------------------------------------------------------------- cut --
def _wedges_bruteforce(nv, tris, cut):
    """Count wedges per vertex by BFS over corners (independent restatement)."""
    cutset = {tuple(sorted(e)) for e in map(tuple, cut)}
    on = np.zeros(nv, bool)
    for u, v in cutset:
        on[u] = on[v] = True
    corners = {}
    for f, t in enumerate(tris):
        for l in range(3):
            corners.setdefault(int(t[l]), []).append(f)
    edge_faces = {}
    for f, t in enumerate(tris):
        for l in range(3):
            e = tuple(sorted((int(t[l]), int(t[(l + 1) % 3]))))
            edge_faces.setdefault(e, []).append(f)
    extra = 0
    for v in range(nv):
        if not on[v]:
            continue
        fs = corners.get(v, [])
        seen, comps = set(), 0
        for f0 in fs:
            if f0 in seen:
                continue
            comps += 1
            stack = [f0]
            seen.add(f0)
            while stack:
                f = stack.pop()
                for w in tris[f]:
                    w = int(w)
                    if w == v:
                        continue
                    e = tuple(sorted((v, w)))
                    if e in cutset:
                        continue
                    for g in edge_faces[e]:
                        if g not in seen and v in tris[g]:
                            seen.add(g)
                            stack.append(g)
        extra += max(comps - 1, 0)
    return extra


def test_cut_cylinder_closed_form():
    m = G.open_cylinder(64, 64)
    nt, origin, con, nv2 = O.cut_edges(m.nv, m.tris, m.cut_edges)
    assert nv2 == 4096 + 64
    assert np.array_equal(origin[4096:], 32 * 64 + np.arange(64))
    assert con.sum() == 128 and con[32 * 64:33 * 64].all() and con[4096:].all()
    # cut mesh = two open cylinders: chi = 0, 4 boundary loops
    h = O.he_from_triangles(nv2, nt)
    assert O.he_manifold(h) == 15
    nb = int((h["face"] < 0).sum())
    assert nb == 4 * 64
    E = (3 * len(nt) + nb) // 2
    assert nv2 - E + len(nt) == 0


def test_cut_sphere_loops_invariants():
    m = G.config_mesh("icosphere6")
    nt, origin, con, nv2 = O.cut_edges(m.nv, m.tris, m.cut_edges)
    assert nv2 - m.nv == _wedges_bruteforce(m.nv, m.tris, m.cut_edges)
    h = O.he_from_triangles(nv2, nt)
    assert O.he_manifold(h) == 15
    # each cut edge becomes two boundary edges
    assert int((h["face"] < 0).sum()) == 2 * len(m.cut_edges)
    # positions of copies are their originals; every copy is constrained
    assert np.all(origin[m.nv:] < m.nv) and con[m.nv:].all()


def test_cut_rejects_non_edges():
    m = G.open_cylinder(8, 8)
    with pytest.raises(ValueError):
        O.cut_edges(m.nv, m.tris, np.array([[0, 20]], np.int32))


# ----------------------------------------------------------- Laplacian CSR --
def _cut_patch(seed):
    nv, tris, pos = grid_patch(7, seed)
    cut = np.array([(3 * 7 + c, 3 * 7 + c + 1) for c in range(1, 5)], np.int32)
    nt, origin, con, nv2 = O.cut_edges(nv, tris, cut)
    return nv2, nt, pos[origin], con


@pytest.mark.parametrize("seed", range(4))
def test_csr_equals_dense_spec_bitwise(seed):
    nv2, nt, pos, con = _cut_patch(seed)
    h = O.he_from_triangles(nv2, nt, spec=True)
    D = O.assemble_vtlhv_dense(h, con, pos)
    r = O.cut_laplacian_csr(h, con, pos)
    free = np.flatnonzero(con == 0)
    cons = np.flatnonzero(con == 1)
    p, c, v = r["A_ff"]
    A = np.zeros((len(free), len(free)))
    for i in range(len(free)):
        A[i, c[p[i]:p[i + 1]]] = v[p[i]:p[i + 1]]
    assert np.array_equal(A, D[np.ix_(free, free)])
    p, c, v = r["A_fc"]
    B = np.zeros((len(free), len(cons)))
    for i in range(len(free)):
        B[i, c[p[i]:p[i + 1]]] = v[p[i]:p[i + 1]]
    # A_fc is the spec's Vt Lh C (the RHS operator, HarmonicSolve.lean:57-64):
    # with Fc = I, -rhs is A_fc entry for entry (products with 1.0 are exact)
    col = np.where(con == 1, np.cumsum(con) - 1, -1).astype(np.int32)
    VtLhC = -O.harmonic_rhs_dense(h, col, pos, np.eye(len(cons)))
    assert np.array_equal(B, VtLhC[free])
    # sample rows / columns of VtLhV are zero (HarmonicSolve.lean:14-16)
    assert np.all(D[cons] == 0) and np.all(D[:, cons] == 0)


@pytest.mark.parametrize("seed", range(3))
def test_harmonic_rhs_csr_equals_dense_spec_bitwise(seed):
    nv2, nt, pos, con = _cut_patch(seed)
    h = O.he_from_triangles(nv2, nt, spec=True)
    r = O.cut_laplacian_csr(h, con, pos)
    col = np.where(con == 1, np.cumsum(con) - 1, -1).astype(np.int32)
    Fc = np.random.default_rng(seed).uniform(-1, 1, (int(con.sum()), 3))
    D = O.harmonic_rhs_dense(h, col, pos, Fc)
    S = O.harmonic_rhs_csr(h, r["free_of"], r["con_of"], pos, Fc)
    assert np.array_equal(S, D[con == 0])


def test_laplacian_symmetric_rowsum_psd():
    m = G.open_cylinder(16, 16)
    h = O.he_from_triangles(m.nv, m.tris)
    r = O.cut_laplacian_csr(h, np.zeros(m.nv, np.uint8), m.pos)
    p, c, v = r["A_ff"]
    n = m.nv
    A = np.zeros((n, n))
    for i in range(n):
        A[i, c[p[i]:p[i + 1]]] = v[p[i]:p[i + 1]]
    assert np.allclose(A, A.T, atol=1e-14)
    assert np.all(np.abs(A.sum(1)) < 1e-12)
    assert np.linalg.eigvalsh(A).min() > -1e-10


def _linear_precision_error(tri_lap):
    """Planar jittered patch, boundary vertices constrained to an affine
    field; the harmonic interior must reproduce it (cot weights have linear
    precision)."""
    n = 8
    nv, tris, pos = grid_patch(n, 3, jit=0.25, shuffle=False)
    bd = np.zeros(nv, bool)
    for i in range(n):
        bd[i] = bd[(n - 1) * n + i] = bd[i * n] = bd[i * n + n - 1] = True
    f = 0.3 + 1.7 * pos[:, 0] - 0.9 * pos[:, 1]
    A = np.zeros((nv, nv))
    for t in tris:
        M = tri_lap(pos[t[1]], pos[t[2]], pos[t[0]])
        loc = [t[1], t[2], t[0]]
        for a in range(3):
            for b in range(3):
                A[loc[a], loc[b]] += M[a, b]
    fr = ~bd
    x = np.linalg.solve(A[np.ix_(fr, fr)], -A[np.ix_(fr, bd)] @ f[bd])
    return np.abs(x - f[fr]).max()


def test_linear_precision():
    assert _linear_precision_error(O.triangle_cot_laplacian) < 1e-10


def test_planted_mistakes_are_caught():
    """Two plausible restatement errors of triangleCotLaplacian fail the
    linear-precision pin: swapped weights and a sign slip."""
    def swapped(a, b, c):
        M = O.triangle_cot_laplacian(a, b, c)
        w_ab, w_ac = -M[0, 1], -M[0, 2]
        M = M.copy()
        M[0, 1] = M[1, 0] = -w_ac
        M[0, 2] = M[2, 0] = -w_ab
        return M

    def sign_slip(a, b, c):
        M = O.triangle_cot_laplacian(a, b, c).copy()
        M[1, 2] = M[2, 1] = -M[1, 2]
        return M

    assert _linear_precision_error(swapped) > 1e-3
    assert _linear_precision_error(sign_slip) > 1e-3


# ------------------------------------------------------------------ IC(0) --
def _csr_dense(p, c, v, n):
    A = np.zeros((n, n))
    for i in range(n):
        A[i, c[p[i]:p[i + 1]]] = v[p[i]:p[i + 1]]
    return A


def _cyl_system(n=16):
    m = G.open_cylinder(n, n, seed=5)
    nt, origin, con, nv2 = O.cut_edges(m.nv, m.tris, m.cut_edges)
    h = O.he_from_triangles(nv2, nt)
    r = O.cut_laplacian_csr(h, con, m.pos[origin])
    return r


def test_ic0_identity_on_pattern():
    """(L L^T)_ij == A_ij on the kept pattern (the defining identity)."""
    r = _cyl_system(16)
    p, c, v = r["A_ff"]
    n = r["n_free"]
    lp, lc, lv = O.ic0_factor(p, c, v)
    L = _csr_dense(lp, lc, lv, n)
    A = _csr_dense(p, c, v, n)
    LLt = L @ L.T
    mask = A != 0
    assert np.max(np.abs(LLt - A)[mask]) <= 1e-12 * np.abs(A).max()
    assert np.all(np.triu(L, 1) == 0)


def test_ic0_full_pattern_is_cholesky():
    rng = np.random.default_rng(2)
    M = rng.standard_normal((7, 7))
    A = M @ M.T + 7 * np.eye(7)
    from kat_data import dense_to_csr
    lp, lc, lv = O.ic0_factor(*dense_to_csr(A, keep=np.ones((7, 7), bool)))
    L = _csr_dense(lp, lc, lv, 7)
    assert np.allclose(L, np.linalg.cholesky(A), rtol=0, atol=1e-13)
    Ld = O.cholesky_factor(A)
    assert np.allclose(Ld, np.linalg.cholesky(A), rtol=0, atol=1e-13)


def test_ic0_solve_and_pcg():
    r = _cyl_system(16)
    p, c, v = r["A_ff"]
    n = r["n_free"]
    lp, lc, lv = O.ic0_factor(p, c, v)
    L = _csr_dense(lp, lc, lv, n)
    b = np.random.default_rng(1).standard_normal(n)
    x = O.ic0_solve(lp, lc, lv, b)
    assert np.allclose(L @ (L.T @ x), b, rtol=0, atol=1e-10)
    A = _csr_dense(p, c, v, n)
    xs = np.linalg.solve(A, b)
    for fac in (None, (lp, lc, lv)):
        xp, it = O.pcg(p, c, v, b, 1e-13 * np.linalg.norm(b), factor=fac)
        assert np.linalg.norm(xp - xs) <= 1e-9 * np.linalg.norm(xs)
    # IC(0) needs fewer iterations than Jacobi
    assert O.pcg(p, c, v, b, 1e-10, factor=(lp, lc, lv))[1] < O.pcg(p, c, v, b, 1e-10)[1]


def test_permute_csr_is_similarity():
    r = _cyl_system(8)
    p, c, v = r["A_ff"]
    n = r["n_free"]
    perm = np.random.default_rng(0).permutation(n).astype(np.int32)
    op, oc, ov = O.permute_csr(p, c, v, perm)
    A = _csr_dense(p, c, v, n)
    assert np.array_equal(_csr_dense(op, oc, ov, n), A[np.ix_(perm, perm)])
    assert all(np.all(np.diff(oc[op[i]:op[i + 1]]) > 0) for i in range(n))


# --------------------------------------------------------------- colouring --
def test_greedy_color_smallest_free():
    r = _cyl_system(16)
    p, c, _ = r["A_ff"]
    color, nc = O.greedy_color(p, c)
    for i in range(len(p) - 1):
        nb = [j for j in c[p[i]:p[i + 1]] if j != i]
        assert all(color[i] != color[j] for j in nb)
        earlier = {int(color[j]) for j in nb if j < i}
        assert color[i] == min(set(range(len(nb) + 2)) - earlier)
    assert 3 <= nc <= 8


# --------------------------------------------------------------------- HSC --
def test_hsc_elimination_is_schur_complement():
    rng = np.random.default_rng(4)
    nv = 7
    edges = [(u, w, float(rng.uniform(0.5, 2.0))) for u in range(nv) for w in range(u + 1, nv)
             if rng.random() < 0.5]
    Lg = np.zeros((nv, nv))
    for u, w, x in edges:
        Lg[u, w] -= x; Lg[w, u] -= x; Lg[u, u] += x; Lg[w, w] += x
    v = 3
    g = O.hsc_eliminate_vertex(edges, v)
    keep = [i for i in range(nv) if i != v]
    S = Lg[np.ix_(keep, keep)] - np.outer(Lg[keep, v], Lg[v, keep]) / Lg[v, v]
    for a in range(len(keep)):
        for b in range(a + 1, len(keep)):
            assert abs(O.hsc_edge_weight(g, keep[a], keep[b]) + S[a, b]) < 1e-12


# ----------------------------------------------------------- dense solves --
def test_dense_against_numpy():
    rng = np.random.default_rng(9)
    A = rng.standard_normal((9, 9)) + 9 * np.eye(9)
    B = rng.standard_normal((9, 4))
    X = O.dense_solve_multi(A, B)
    assert np.allclose(X, np.linalg.solve(A, B), atol=1e-12)
    a, piv = O.lu_factor(A)
    assert np.allclose(O.lu_solve(a, piv, B[:, 0]), np.linalg.solve(A, B[:, 0]), atol=1e-12)


# ------------------------------------------------------- knot projection --
def _first_argmin(knots, pos):
    """numpy restatement: the lowest index among the minima of the squared
    distances (np.argmin returns the first minimum)."""
    d = pos[None, :, :] - knots[:, None, :]
    dd = d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1] + d[..., 2] * d[..., 2]
    return np.argmin(dd, axis=1).astype(np.int32)


def _cloud_with_ties(seed):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-1, 1, (60, 3))
    pos[40:50] = pos[5:15]                  # duplicated vertices: exact distance ties
    knots = np.concatenate([rng.uniform(-1, 1, (40, 3)), pos[5:15] + 0.0])
    # knots equidistant from two vertices: the midpoints of vertex pairs on
    # an axis-aligned grid, where both squared distances round identically
    g = np.array([[0.0, 0.0, 0.0], [0.5, 0.0, 0.0], [0.0, 0.5, 0.0], [0.5, 0.5, 0.0]])
    pos = np.concatenate([pos, g + 3.0])
    knots = np.concatenate([knots, [[3.25, 3.0, 3.0], [3.0, 3.25, 3.0], [3.25, 3.25, 3.0]]])
    return knots, pos


@pytest.mark.parametrize("seed", range(3))
def test_projection_first_argmin(seed):
    knots, pos = _cloud_with_ties(seed)
    got = O.project_to_vertices(knots, pos)
    assert np.array_equal(got, _first_argmin(knots, pos))
    # the duplicated vertices and the equidistant knots resolve to the lower index
    assert (got[40:50] == np.arange(5, 15)).all()
    assert list(got[-3:]) == [60, 60, 60]


def test_projection_planted_last_index_fails():
    knots, pos = _cloud_with_ties(0)
    # a "<=" scan (last minimum wins) is the plausible mistake
    d = pos[None] - knots[:, None]
    dd = (d * d).sum(-1)
    last = (dd.shape[1] - 1 - np.argmin(dd[:, ::-1], axis=1)).astype(np.int32)
    assert not np.array_equal(O.project_to_vertices(knots, pos), last)


def _promote_closed_form(nv, idx):
    """Columns = rank of each vertex's first occurrence among the valid knots."""
    idx = np.asarray(idx)
    valid = np.flatnonzero(idx >= 0)
    kc = np.full(idx.shape[0], -1, np.int32)
    vc = np.full(nv, -1, np.int32)
    if valid.size:
        verts, first = np.unique(idx[valid], return_index=True)
        order = np.argsort(valid[first], kind="stable")       # encounter order
        vc[verts[order]] = np.arange(order.size)
        kc[valid] = vc[idx[valid]]
    return kc, vc, int((vc >= 0).sum())


@pytest.mark.parametrize("seed", range(4))
def test_promote_closed_form(seed):
    rng = np.random.default_rng(seed)
    nv = 50
    idx = rng.integers(-1, nv, 300).astype(np.int32)
    kc, vc, nc = O.promote_vertex_samples(nv, idx)
    ekc, evc, enc = _promote_closed_form(nv, idx)
    assert np.array_equal(kc, ekc) and np.array_equal(vc, evc) and nc == enc


def test_promote_planted_vertex_order_fails():
    idx = np.array([7, 3, 7, 1, 3], np.int32)
    kc, _, _ = O.promote_vertex_samples(8, idx)
    assert list(kc) == [0, 1, 0, 2, 1]
    # columns numbered by vertex index instead of encounter order
    wrong = np.searchsorted(np.unique(idx), idx)
    assert not np.array_equal(kc, wrong)


def test_promote_rejects_out_of_range():
    with pytest.raises(IndexError):
        O.promote_vertex_samples(3, [0, 3])


def test_every_oracle_function_is_pinned():
    import inspect
    import re
    import oracle.oracle as OO
    tokens = set()
    for key in OO.PINS:
        tokens |= set(re.findall(r"[A-Za-z0-9_]+", key))
    infra = {"build", "lib"}
    fns = [n for n, f in inspect.getmembers(OO, inspect.isfunction)
           if not n.startswith("_") and f.__module__ == OO.__name__ and n not in infra]
    assert [f for f in fns if f not in tokens] == []
