"""Pins of the oracle's adjoint and of the B-spline basis (CPU only).

The adjoint is checked against central finite differences of the loss
L = 1/2 |x_f - delta|^2 with x_f = A_ff^-1 (-A_fc x_c), computed with a
dense FP64 solve (numpy), both in the constrained values x_c and in edge
weights w (A_ii += w, A_jj += w, A_ij = A_ji -= w)."""
import numpy as np
import pytest

import oracle as O
from tool_godot_curvenet_amd import curves, meshgen


@pytest.fixture(scope="module")
def small():
    m = meshgen.open_cylinder(12, 10, seed=7)
    nt, origin, con, nv2 = O.cut_edges(m.nv, m.tris, m.cut_edges)
    pos = m.pos[origin]
    he = O.he_from_triangles(nv2, nt)
    r = O.cut_laplacian_csr(he, con, pos)
    n, ncon = r["n_free"], int(con.sum())
    ff = r["A_ff"]
    fc = r["A_fc"]
    Aff = np.zeros((n, n))
    Afc = np.zeros((n, ncon))
    for i in range(n):
        for e in range(ff[0][i], ff[0][i + 1]):
            Aff[i, ff[1][e]] = ff[2][e]
        for e in range(fc[0][i], fc[0][i + 1]):
            Afc[i, fc[1][e]] = fc[2][e]
    rng = np.random.default_rng(3)
    xc = rng.uniform(-0.2, 0.2, size=(ncon, 2))
    delta = rng.normal(0, 0.05, size=(n, 2))
    return dict(ff=ff, fc=(fc[0], fc[1], fc[2], ncon), Aff=Aff, Afc=Afc, xc=xc, delta=delta,
                n=n, ncon=ncon)


def _loss(Aff, Afc, xc, delta):
    x = np.linalg.solve(Aff, -Afc @ xc)
    return 0.5 * ((x - delta) ** 2).sum(), x


def _grads(S, sign=1.0, transpose_ok=True):
    L0, x = _loss(S["Aff"], S["Afc"], S["xc"], S["delta"])
    lam, gc, gw = O.adjoint(S["ff"], S["fc"], x - S["delta"], x=x, xc=S["xc"])
    return x, lam, gc, gw


def test_adjoint_lambda_solves(small):
    S = small
    x, lam, gc, gw = _grads(S)
    assert np.allclose(S["Aff"] @ lam, x - S["delta"], rtol=0, atol=1e-11)


def test_adjoint_gc_matches_fd(small):
    S = small
    x, lam, gc, gw = _grads(S)
    rng = np.random.default_rng(5)
    h = 1e-6
    for _ in range(16):
        c, j = rng.integers(S["ncon"]), rng.integers(2)
        xp, xm = S["xc"].copy(), S["xc"].copy()
        xp[c, j] += h
        xm[c, j] -= h
        fd = (_loss(S["Aff"], S["Afc"], xp, S["delta"])[0]
              - _loss(S["Aff"], S["Afc"], xm, S["delta"])[0]) / (2 * h)
        assert abs(gc[c, j] - fd) <= 1e-6 * max(1.0, abs(fd)), (c, j, gc[c, j], fd)


def test_adjoint_gw_matches_fd(small):
    S = small
    x, lam, gc, gw = _grads(S)
    ptr, col, _ = S["ff"]
    fptr, fcol, _, _ = S["fc"]
    n = S["n"]
    rng = np.random.default_rng(6)
    h = 1e-6
    off = [(i, e) for i in range(n) for e in range(ptr[i], ptr[i + 1]) if col[e] != i]
    for idx in rng.choice(len(off), 8, replace=False):
        i, e = off[idx]
        j = col[e]
        Ls = []
        for sgn in (1, -1):
            A = S["Aff"].copy()
            A[i, i] += sgn * h
            A[j, j] += sgn * h
            A[i, j] -= sgn * h
            A[j, i] -= sgn * h
            Ls.append(_loss(A, S["Afc"], S["xc"], S["delta"])[0])
        fd = (Ls[0] - Ls[1]) / (2 * h)
        assert abs(gw[e] - fd) <= 1e-6 * max(1.0, abs(fd)), (i, j, gw[e], fd)
    nnz = len(col)
    fc_e = [(i, e) for i in range(n) for e in range(fptr[i], fptr[i + 1])]
    for idx in rng.choice(len(fc_e), 8, replace=False):
        i, e = fc_e[idx]
        c = fcol[e]
        Ls = []
        for sgn in (1, -1):
            A = S["Aff"].copy()
            B = S["Afc"].copy()
            A[i, i] += sgn * h
            B[i, c] -= sgn * h
            Ls.append(_loss(A, B, S["xc"], S["delta"])[0])
        fd = (Ls[0] - Ls[1]) / (2 * h)
        assert abs(gw[nnz + e] - fd) <= 1e-6 * max(1.0, abs(fd)), (i, c, gw[nnz + e], fd)


def test_adjoint_planted_mistakes_fail_fd(small):
    """A dropped minus sign or a missing transpose in gc would not pass."""
    S = small
    x, lam, gc, gw = _grads(S)
    h = 1e-6
    c, j = 0, 0
    xp, xm = S["xc"].copy(), S["xc"].copy()
    xp[c, j] += h
    xm[c, j] -= h
    fd = (_loss(S["Aff"], S["Afc"], xp, S["delta"])[0]
          - _loss(S["Aff"], S["Afc"], xm, S["delta"])[0]) / (2 * h)
    wrong_sign = -gc[c, j]
    wrong_op = -(S["Afc"].T @ np.linalg.solve(S["Aff"], S["delta"]))[c, j]   # forgot x
    assert abs(gc[c, j] - fd) < 1e-6 * max(1.0, abs(fd))
    assert abs(wrong_sign - fd) > 1e-3 * abs(fd)
    assert abs(wrong_op - fd) > 1e-3 * abs(fd)


# ------------------------------------------------------------- B-splines --
def test_bspline_knot_values_and_partition():
    n = 16
    idx, w = curves.bspline_weights(np.arange(n) / n, n)
    assert np.allclose(w, [[1 / 6, 4 / 6, 1 / 6, 0.0]] * n, atol=1e-15)
    assert np.all(idx[:, 1] == np.arange(n))
    t = np.random.default_rng(1).uniform(0, 1, 200)
    _, w = curves.bspline_weights(t, n)
    assert np.allclose(w.sum(1), 1.0, atol=1e-14) and np.all(w >= 0)


def _cox_de_boor(j, p, u):
    """B-spline basis N_{j,p}(u) on the integer knots (Cox-de Boor)."""
    if p == 0:
        return 1.0 if j <= u < j + 1 else 0.0
    return ((u - j) / p * _cox_de_boor(j, p - 1, u)
            + (j + p + 1 - u) / p * _cox_de_boor(j + 1, p - 1, u))


def test_bspline_matches_cox_de_boor():
    """Closed uniform cubic weights equal the Cox-de Boor cubic basis on a
    periodic integer knot vector: control point q at parameter u = n t has
    weight sum over periods N_{q-1+m n, 3}(u)."""
    n = 8
    rng = np.random.default_rng(4)
    for t in rng.uniform(0, 1, 40):
        idx, w = curves.bspline_weights(np.array([t]), n)
        u = n * t
        ref = np.zeros(n)
        for q in range(n):
            for m in (-1, 0, 1):
                ref[q] += _cox_de_boor(q - 2 + m * n, 3, u)
        got = np.zeros(n)
        np.add.at(got, idx[0], w[0])
        assert np.allclose(got, ref, atol=1e-14), (t, got, ref)


def test_constrained_basis_rows_sum_to_one():
    m = meshgen.config_mesh("sphere45")
    nt, origin, con, nv2 = O.cut_edges(m.nv, m.tris, m.cut_edges)
    c64 = con.astype(np.int64)
    con_of = np.where(c64 > 0, np.cumsum(c64) - 1, -1)
    Bm = curves.constrained_basis_from(origin, con_of, int(con.sum()), m, 16)
    assert Bm.shape == (int(con.sum()), 8 * 16)
    assert np.allclose(Bm.sum(1), 1.0, atol=1e-14)
    assert np.all((Bm > 0).sum(1) <= 4)
