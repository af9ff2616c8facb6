"""Pins of the oracle's colour-parallel symmetric Gauss-Seidel and of the
coarse-vertex restriction (CPU only).

SGS: one symmetric sweep from x = 0 is the textbook SGS preconditioner
z = M^-1 r with M = (D + L) D^-1 (D + U), L / U the strictly lower / upper
parts of A in the sweep's (colour) order -- a matrix identity independent of
the row-update loop the oracle restates; and the sweep's fixed point is the
solution of A x = b."""
import numpy as np
import pytest

import oracle as O
from tool_godot_curvenet_amd import meshgen


def _dense(ptr, col, val, n):
    A = np.zeros((n, n))
    for i in range(n):
        A[i, col[ptr[i]:ptr[i + 1]]] = val[ptr[i]:ptr[i + 1]]
    return A


@pytest.fixture(scope="module")
def cyl():
    m = meshgen.open_cylinder(12, 10, seed=5)
    nt, origin, con, nv2 = O.cut_edges(m.nv, m.tris, m.cut_edges)
    he = O.he_from_triangles(nv2, nt)
    ptr, col, val = O.cut_laplacian_csr(he, con, m.pos[origin])["A_ff"]
    return ptr, col, val, _dense(ptr, col, val, len(ptr) - 1)


def _sgs_matrix(A, order, swap=False):
    """(D+L) D^-1 (D+U) with L/U taken in `order` (swap: the reverse
    product, a plausible mistake)."""
    P = A[np.ix_(order, order)]
    D = np.diag(np.diag(P))
    L, U = np.tril(P, -1), np.triu(P, 1)
    M = (D + U) @ np.linalg.inv(D) @ (D + L) if swap else (D + L) @ np.linalg.inv(D) @ (D + U)
    Mi = np.zeros_like(M)
    Mi[np.ix_(order, order)] = M
    return Mi


def test_colouring_is_valid(cyl):
    ptr, col, val, A = cyl
    color, nc = O.greedy_color(ptr, col)
    for i in range(len(A)):
        nb = col[ptr[i]:ptr[i + 1]]
        assert not np.any(color[nb[nb != i]] == color[i])
    assert 3 <= nc <= 8


def test_one_sweep_from_zero_is_the_sgs_preconditioner(cyl):
    ptr, col, val, A = cyl
    color, nc = O.greedy_color(ptr, col)
    _, order = O.color_classes(color, nc)
    r = np.random.default_rng(1).normal(size=(len(A), 3))
    z = O.sgs_apply(ptr, col, val, r)
    M = _sgs_matrix(A, order)
    assert np.allclose(M @ z, r, rtol=0, atol=1e-12 * np.abs(r).max())
    assert np.allclose(M, M.T, atol=1e-12)                  # symmetric preconditioner
    # the reverse product (backward sweep first) is a different operator
    assert not np.allclose(_sgs_matrix(A, order, swap=True) @ z, r, atol=1e-6)


def test_sweeps_converge_to_the_solution(cyl):
    ptr, col, val, A = cyl
    b = np.random.default_rng(2).normal(size=len(A))
    x = O.sgs_sweep(ptr, col, val, b, np.zeros(len(A)), sweeps=400)
    assert np.allclose(x, np.linalg.solve(A, b), rtol=0, atol=1e-9 * np.abs(x).max())


def test_zero_diagonal_row_is_left_alone():
    # [[2,1,0],[1,0,1],[0,1,2]]: row 1 has A_11 = 0 (SgsColor.lean:107-108)
    ptr = np.array([0, 2, 5, 7], np.int32)
    col = np.array([0, 1, 0, 1, 2, 1, 2], np.int32)
    val = np.array([2.0, 1.0, 1.0, 0.0, 1.0, 1.0, 2.0])
    x0 = np.array([0.5, 7.0, -0.5])
    x = O.sgs_sweep(ptr, col, val, np.array([1.0, 1.0, 1.0]), x0, sweeps=3)
    assert x[1] == 7.0
    assert x[0] == (1.0 - 7.0) / 2.0 and x[2] == (1.0 - 7.0) / 2.0


def test_fp32_sweep_tracks_fp64(cyl):
    ptr, col, val, A = cyl
    b = np.random.default_rng(3).normal(size=(len(A), 2))
    x64 = O.sgs_sweep(ptr, col, val, b, np.zeros_like(b), sweeps=2)
    x32 = O.sgs_sweep(ptr, col, val, b, np.zeros_like(b), sweeps=2, dtype=np.float32)
    assert x32.dtype == np.float32
    assert np.abs(x32 - x64).max() <= 1e-5 * np.abs(x64).max()


def test_restrict_sum_kat_and_closed_form():
    # parent -1 and parents >= n_coarse match no coarse vertex
    parent = np.array([0, 1, 0, 2, 1, -1, 3], np.int32)
    fine = np.array([1.0, 2.0, 3.0, 4.0, 5.0, 6.0, 7.0])
    assert np.array_equal(O.restrict_sum(parent, 3, fine), [4.0, 7.0, 4.0])
    rng = np.random.default_rng(4)
    parent = rng.integers(-1, 40, size=500).astype(np.int32)
    fine = rng.normal(size=(500, 5))
    want = np.zeros((40, 5))
    ok = parent >= 0
    np.add.at(want, parent[ok], fine[ok])       # in index order: the same sums
    assert np.array_equal(O.restrict_sum(parent, 40, fine), want)
    f32 = O.restrict_sum(parent, 40, fine, dtype=np.float32)
    want32 = np.zeros((40, 5), np.float32)
    np.add.at(want32, parent[ok], fine[ok].astype(np.float32))
    assert np.array_equal(f32, want32)


def test_pcg_many_equals_single_column_solves(cyl):
    """The threaded CPU baseline solves every column with exactly the
    single-column PCG's arithmetic."""
    ptr, col, val, A = cyl
    B = np.random.default_rng(6).normal(size=(len(A), 5))
    fac = O.ic0_factor(ptr, col, val)
    X, it = O.pcg_many(ptr, col, val, B, 1e-10, factor=fac, threads=3)
    total = 0
    for c in range(5):
        bb = 0.0
        for v in B[:, c]:
            bb = bb + v * v
        x, i = O.pcg(ptr, col, val, B[:, c], 1e-10 * bb ** 0.5, factor=fac)
        assert np.array_equal(X[:, c], x)
        total += i
    assert it == total
