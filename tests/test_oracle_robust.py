"""Pins of the oracle's robust (Sharp-Crane) triangle Laplacian, the
RobustLaplacian.lean:160-290 known answers plus a closed-form cross-check:
with delta = 0 on a non-degenerate triangle the law-of-cosines/Heron cot
equals the embedded cross-product cot (a different formula)."""
import math

import numpy as np

import oracle as O

EQ = (np.array([0.0, 0.0, 0.0]), np.array([1.0, 0.0, 0.0]),
      np.array([0.5, 0.0, 0.8660254037844386]))


def test_mollify_kats():
    assert abs(O.mollify_triangle(0.0, 1.0, 1.0, 1.0)) < 1e-12
    assert abs(O.mollify_triangle(0.01, 1.0, 1.0, 1.0)) < 1e-12
    eps = O.mollify_triangle(0.05, 1.0, 1.0, 1.99)
    assert eps > 0.0 and abs(eps - (0.05 + 1.99 - 2.0)) < 1e-15
    e = np.array([1.0, 1.0, 1.99]) + eps
    tol = 1e-12
    assert e[1] + e[2] > e[0] + 0.05 - tol
    assert e[0] + e[2] > e[1] + 0.05 - tol
    assert e[0] + e[1] > e[2] + 0.05 - tol


def test_cot_from_lengths_kats():
    assert abs(O.cot_from_lengths(1.0, 1.0, 1.0) - 1.0 / math.sqrt(3.0)) < 1e-12
    assert abs(O.cot_from_lengths(1.0, math.sqrt(3.0), 2.0) - math.sqrt(3.0)) < 1e-9
    assert abs(O.cot_from_lengths(math.sqrt(3.0), 1.0, 2.0) - 1.0 / math.sqrt(3.0)) < 1e-9
    assert abs(O.cot_from_lengths(2.0, 1.0, math.sqrt(3.0))) < 1e-9


def test_equilateral_equals_embedded_and_is_a_laplacian():
    L = O.robust_triangle_cot_laplacian(0.0, *EQ)
    M = np.zeros(9)
    import ctypes as C
    O.lib().orc_triangle_cot_laplacian(*(EQ[i].ctypes.data_as(C.c_void_p) for i in range(3)),
                                       M.ctypes.data_as(C.c_void_p))
    assert np.abs(L - M.reshape(3, 3)).max() < 1e-12
    assert np.abs(L.sum(1)).max() < 1e-12 and np.abs(L - L.T).max() < 1e-12
    assert abs(L[0, 1] + 1.0 / (2 * math.sqrt(3.0))) < 1e-12


def test_collinear_triangle_is_finite():
    L = O.robust_triangle_cot_laplacian(0.01, np.zeros(3), np.array([1.0, 0, 0]),
                                        np.array([1.999999, 0, 0]))
    assert np.all(np.isfinite(L))
    assert np.abs(L.sum(1)).max() < 1e-9


def test_random_triangles_match_embedded_cot():
    import ctypes as C
    rng = np.random.default_rng(3)
    for _ in range(200):
        a, b, c = rng.normal(size=(3, 3))
        L = O.robust_triangle_cot_laplacian(0.0, a, b, c)
        M = np.zeros(9)
        O.lib().orc_triangle_cot_laplacian(a.ctypes.data_as(C.c_void_p),
                                           b.ctypes.data_as(C.c_void_p),
                                           c.ctypes.data_as(C.c_void_p),
                                           M.ctypes.data_as(C.c_void_p))
        # skip slivers, where both forms lose digits
        if np.abs(M).max() > 50:
            continue
        assert np.abs(L - M.reshape(3, 3)).max() <= 1e-9 * max(1.0, np.abs(M).max())


def test_planted_corner_mistake_fails():
    """Taking the requirement of one corner instead of the max over the
    three leaves the collinear triangle degenerate."""
    e = (1.0, 1.0, 1.999999)
    wrong = max(0.0, 0.01 + e[0] - e[1] - e[2])      # corner a only
    ee = np.array(e) + wrong
    assert not (ee[0] + ee[1] > ee[2] + 0.01)
    ok = np.array(e) + O.mollify_triangle(0.01, *e)
    assert ok[0] + ok[1] >= ok[2] + 0.01 - 1e-12
