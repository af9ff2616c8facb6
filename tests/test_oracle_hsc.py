"""Pins of the oracle's hierarchical Schur preconditioner (CPU only).

With theta = 0 nothing is dropped, so every level is the exact Schur
complement (textbook: S = A_CC - A_CF A_FF^-1 A_CF^T) and the hierarchy is an
exact block factorisation: its apply must equal A^-1.  With theta > 0 the
compensation rules are checked against their stated invariants."""
import numpy as np
import pytest

import oracle as O
from kat_data import csr_np
from tool_godot_curvenet_amd import meshgen


def _dense(ptr, col, val, n):
    A = np.zeros((n, n))
    for i in range(n):
        A[i, col[ptr[i]:ptr[i + 1]]] = val[ptr[i]:ptr[i + 1]]
    return A


@pytest.fixture(scope="module")
def cyl():
    m = meshgen.open_cylinder(16, 12, seed=9)
    nt, origin, con, nv2 = O.cut_edges(m.nv, m.tris, m.cut_edges)
    he = O.he_from_triangles(nv2, nt)
    r = O.cut_laplacian_csr(he, con, m.pos[origin])
    ptr, col, val = r["A_ff"]
    return ptr, col, val, _dense(ptr, col, val, r["n_free"])


def test_one_level_is_the_dense_schur_complement(cyl):
    ptr, col, val, A = cyl
    n = len(A)
    h = O.hsc_build(ptr, col, val, theta=0.0, coarse_max=n - 1, max_levels=1)
    L = h["levels"][0]
    F, Cc = L["F"], L["C"]
    S = A[np.ix_(Cc, Cc)] - A[np.ix_(Cc, F)] @ np.linalg.solve(A[np.ix_(F, F)], A[np.ix_(F, Cc)])
    assert np.allclose(h["coarse"], S, rtol=0, atol=1e-13 * np.abs(A).max())
    # F independent and maximal; visited by (off-diagonal count, index)
    isF = np.zeros(n, bool)
    isF[F] = True
    for f in F:
        nb = col[ptr[f]:ptr[f + 1]]
        assert not np.any(isF[nb[nb != f]])
    for c in Cc:
        nb = col[ptr[c]:ptr[c + 1]]
        assert np.any(isF[nb[nb != c]])
    deg = np.diff(ptr) - 1
    first = min(range(n), key=lambda i: (deg[i], i))
    assert isF[first]


@pytest.mark.parametrize("levels", [1, 3, 24])
def test_exact_hierarchy_apply_is_the_inverse(cyl, levels):
    ptr, col, val, A = cyl
    h = O.hsc_build(ptr, col, val, theta=0.0, coarse_max=16, max_levels=levels)
    assert len(h["levels"]) >= 1
    r = np.random.default_rng(1).normal(size=(len(A), 3))
    z = O.hsc_apply(h, r)
    assert np.allclose(A @ z, r, rtol=0, atol=1e-10 * np.abs(r).max())


def test_planted_sign_mistake_fails_exactness(cyl):
    ptr, col, val, A = cyl
    h = O.hsc_build(ptr, col, val, theta=0.0, coarse_max=16, max_levels=2)
    rp, rc, rv = h["levels"][0]["R"]
    h["levels"][0]["R"] = (rp, rc, -rv)
    r = np.random.default_rng(2).normal(size=len(A))
    z = O.hsc_apply(h, r)
    assert not np.allclose(A @ z, r, atol=1e-6)


def test_lump_keeps_row_sums_diag_only_grows(cyl):
    ptr, col, val, A = cyl
    n = len(A)
    ex = O.hsc_build(ptr, col, val, theta=0.0, coarse_max=n - 1, max_levels=1)["coarse"]
    lump = O.hsc_build(ptr, col, val, theta=0.5, coarse_max=n - 1, max_levels=1, comp="lump")["coarse"]
    diag = O.hsc_build(ptr, col, val, theta=0.5, coarse_max=n - 1, max_levels=1, comp="diag")["coarse"]
    assert np.count_nonzero(lump) < np.count_nonzero(ex)       # something was dropped
    assert np.allclose(lump.sum(1), ex.sum(1), atol=1e-12)
    assert np.linalg.eigvalsh(diag - ex).min() >= -1e-12
    assert np.linalg.eigvalsh(diag).min() > 0


@pytest.mark.parametrize("edges,v", [
    ([(0, 1, 1.0), (1, 2, 1.0)], 1),                                   # path3
    ([(0, 1, 1.0), (0, 2, 1.0), (0, 3, 1.0)], 0),                      # star4
    ([(0, 1, 1.0), (0, 2, 1.0), (1, 2, 1.0), (2, 3, 1.0)], 2),          # triangle + leaf
])
def test_level_rule_is_eliminate_vertex(edges, v):
    """HierarchicalSparsifyCompensate.lean:134-231 graphs: the Schur
    complement of the graph Laplacian with respect to v (the operation a
    level applies to its independent set, pinned against the dense formula
    above) has exactly the edge weights of HSC.eliminateVertex."""
    nv = 1 + max(max(a, b) for a, b, _ in edges)
    Lp = np.zeros((nv, nv))
    for a, b, w in edges:
        Lp[a, b] -= w
        Lp[b, a] -= w
        Lp[a, a] += w
        Lp[b, b] += w
    keep = [i for i in range(nv) if i != v]
    S = Lp[np.ix_(keep, keep)] - np.outer(Lp[keep, v], Lp[v, keep]) / Lp[v, v]
    g = O.hsc_eliminate_vertex(edges, v)
    W = np.zeros((nv, nv))
    for a, b, w in g:
        W[a, b] = W[b, a] = w
    for x, i in enumerate(keep):
        for y, j in enumerate(keep):
            if i != j:
                assert abs(-S[x, y] - W[i, j]) < 1e-15, (i, j, -S[x, y], W[i, j])
