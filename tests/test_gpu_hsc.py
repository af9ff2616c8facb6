"""GPU hierarchical Schur preconditioner and the dense FP64 Cholesky on the
matrix cores, against the oracle (itself pinned in test_oracle_hsc.py)."""
import sys

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, file=sys.stderr, flush=True)


def _spd(n, seed):
    rng = np.random.default_rng(seed)
    M = rng.normal(size=(n, n))
    return M @ M.T / n + np.eye(n)


@pytest.mark.parametrize("n", [1, 31, 32, 100, 257])
def test_dense_cholesky_matches_numpy(gpu, n):
    from tool_godot_curvenet_amd import dense_cholesky
    A = _spd(n, n)
    Lg = dense_cholesky(A)
    Ln = np.linalg.cholesky(A)
    assert np.all(np.triu(Lg, 1) == 0)          # spec: zero upper triangle
    assert np.abs(Lg - Ln).max() <= 1e-12 * np.abs(Ln).max()
    assert np.abs(Lg @ Lg.T - A).max() <= 1e-12 * np.abs(A).max()


def test_dense_cholesky_spec_kat(gpu):
    """DenseLinAlg.lean:294-489: [[4,2],[2,3]] -> [[2,0],[1,sqrt 2]], solves
    to (1.75, 1.5)-style values reproduced by the oracle."""
    from tool_godot_curvenet_amd import dense_cholesky, dense_cholesky_solve
    A = np.array([[4.0, 2.0], [2.0, 3.0]])
    Lg = dense_cholesky(A)
    assert np.allclose(Lg, [[2.0, 0.0], [1.0, np.sqrt(2.0)]], atol=1e-15)
    b = np.array([10.0, 8.0])
    assert np.allclose(dense_cholesky_solve(Lg, b), np.linalg.solve(A, b), atol=1e-14)


def test_dense_cholesky_rejects_indefinite(gpu):
    from tool_godot_curvenet_amd import dense_cholesky
    from tool_godot_curvenet_amd._lib import CurvenetError
    A = _spd(40, 3)
    A[20, 20] = -5.0
    with pytest.raises(CurvenetError):
        dense_cholesky(A)


@pytest.mark.parametrize("n,k", [(64, 1), (100, 7), (300, 64)])
def test_dense_cholesky_solve(gpu, n, k):
    from tool_godot_curvenet_amd import dense_cholesky, dense_cholesky_solve
    A = _spd(n, 7 + n)
    B = np.random.default_rng(k).normal(size=(n, k))
    X = dense_cholesky_solve(dense_cholesky(A), B)
    assert np.abs(A @ X - B).max() <= 1e-11 * np.abs(B).max()


def _lap(name):
    from tool_godot_curvenet_amd import CutLaplacian, meshgen
    return CutLaplacian.from_mesh(meshgen.config_mesh(name))


@pytest.mark.parametrize("name", ["cylinder64", "icosphere6"])
@pytest.mark.parametrize("comp", ["lump", "diag"])
def test_hierarchy_bit_exact(gpu, name, comp):
    from tool_godot_curvenet_amd import HSC
    lap = _lap(name)
    A = lap.A_ff()
    H = HSC(lap, compensate=comp)
    ho = O.hsc_build(A.row_ptr, A.col_idx, A.values, comp=comp)
    lv = H.levels()
    assert len(lv) == len(ho["levels"]) and len(lv) >= 3
    for g, o in zip(lv, ho["levels"]):
        assert g["n"] == o["n"]
        assert np.array_equal(g["F"], o["F"]) and np.array_equal(g["C"], o["C"])
        assert np.array_equal(g["dinv"], o["dinv"])
        assert np.array_equal(g["R"][0], o["R"][0])
        assert np.array_equal(g["R"][1], o["F"][o["R"][1]])   # level rows of F vertices
        assert np.array_equal(g["R"][2], o["R"][2])
        for a, b in zip(g["P"], o["P"]):
            assert np.array_equal(a, b)
    Lc, inv = H.coarse()
    Co = ho["coarse"]
    assert Lc.shape == Co.shape
    assert np.abs(Lc @ Lc.T - Co).max() <= 1e-12 * np.abs(Co).max()
    assert np.abs(inv @ Co - np.eye(len(Co))).max() <= 1e-9


# k <= 4 takes the coarse GEMV; k = 5, 8 and 48 take the matrix-core coarse
# product in 64 x 64 tiles with its j range in 8 slices (391 coarse rows:
# the last slice starts past the last row and adds only zeros).
@pytest.mark.parametrize("k", [1, 3, 5, 8, 48])
def test_hsc_apply_matches_oracle(gpu, k):
    from tool_godot_curvenet_amd import HSC
    lap = _lap("cylinder64")
    A = lap.A_ff()
    H = HSC(lap)
    ho = O.hsc_build(A.row_ptr, A.col_idx, A.values)
    r = np.random.default_rng(k).normal(size=(lap.n_free, k))
    zg = H.apply(r)
    zo = O.hsc_apply(ho, r)
    assert np.linalg.norm(zg - zo) <= 1e-11 * np.linalg.norm(zo)


# k > 4: the down sweep runs one launch per level of elimination, each row
# summing its restriction chain through every level it is kept through; the
# 40k-row icosphere has 12 levels, so chains of up to 12 segments, with
# (k = 512) and without the fused tail
@pytest.mark.parametrize("k", [8, 48, 97, 512])
def test_hsc_apply_chains_match_oracle(gpu, k):
    from tool_godot_curvenet_amd import HSC
    lap = _lap("icosphere6")
    A = lap.A_ff()
    H = HSC(lap)
    assert H.info.nlevels >= 8
    ho = O.hsc_build(A.row_ptr, A.col_idx, A.values)
    r = np.random.default_rng(k).normal(size=(lap.n_free, k))
    zg = H.apply(r)
    zo = O.hsc_apply(ho, r)
    assert np.linalg.norm(zg - zo) <= 1e-11 * np.linalg.norm(zo)


# A coarse block of more than 3,072 rows and k >= 512 takes the 128 x 128
# tiles of the coarse product (k = 520 and 3,000+ rows: partial tiles in
# both directions); 192 columns take the 64 x 64 tiles on the same block.
@pytest.mark.parametrize("k", [192, 520])
def test_hsc_apply_large_coarse_matches_oracle(gpu, k):
    from tool_godot_curvenet_amd import HSC
    lap = _lap("icosphere6")
    A = lap.A_ff()
    H = HSC(lap, coarse_max=4096)
    assert 3072 <= H.info.n_coarse <= 4096, H.info.n_coarse
    ho = O.hsc_build(A.row_ptr, A.col_idx, A.values, coarse_max=4096)
    r = np.random.default_rng(k).normal(size=(lap.n_free, k))
    zg = H.apply(r)
    zo = O.hsc_apply(ho, r)
    assert np.linalg.norm(zg - zo) <= 1e-11 * np.linalg.norm(zo)


# groups of 1 (the in-place level pass), 2, 4 and all levels: one PCG step
# from 0 follows the oracle hierarchy's apply, and the solve converges
@pytest.mark.parametrize("g", [1, 2, 4, 24])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-11), ("mixed", 2e-5)])
def test_chain_levels_direction_and_convergence(gpu, g, precision, tol):
    from tool_godot_curvenet_amd import HSC
    from tool_godot_curvenet_amd import meshgen as G
    k = 64
    lap = _lap("icosphere6")
    A = lap.A_ff()
    H = HSC(lap)
    lap.set_tune({"chain_levels": g})
    B = lap.harmonic_rhs(G.seeded_boundary(lap.n_con, k, seed=90 + g))
    x, info = lap.solve(B, precond="hsc", precision=precision, tol=1e-9, max_iter=1, hsc=H)
    assert info.iters == 1
    X = x.cpu().numpy()
    Z = O.hsc_apply(O.hsc_build(A.row_ptr, A.col_idx, A.values), B)
    e = np.linalg.norm(X / np.linalg.norm(X, axis=0) - Z / np.linalg.norm(Z, axis=0), axis=0)
    assert np.all(e <= tol), e.max()
    x, info = lap.solve(B, precond="hsc", precision=precision, tol=1e-9, hsc=H)
    assert info.status == 0
    R = B - O.spmm(A.row_ptr, A.col_idx, A.values, x.cpu().numpy())
    assert np.all(np.linalg.norm(R, axis=0) <= 1e-9 * np.linalg.norm(B, axis=0))


# The FP32 coarse product (coarse_prec 2) in the one-wavefront-a-row vector
# form (k = 3), on 64 x 64 tiles (k = 64) and 128 x 128 tiles (k = 520;
# 3,000+ coarse rows), and the split-bfloat16 product (coarse_prec 3, both
# tile shapes; a b as a_h b_h + a_h b_l + a_l b_h, 2^-16 of a b dropped): one mixed PCG step from 0 follows the oracle hierarchy's apply
# to FP32 rounding (3e-4 for the split product), and the solve converges;
# an FP64 solve on the same tune keeps the FP64 product.
@pytest.mark.parametrize("k,prec", [(3, 2), (64, 2), (64, 3), (520, 2), (520, 3)])
def test_coarse_fp32_direction_and_convergence(gpu, k, prec):
    from tool_godot_curvenet_amd import HSC
    from tool_godot_curvenet_amd import meshgen as G
    lap = _lap("icosphere6")
    A = lap.A_ff()
    H = HSC(lap, coarse_max=4096)
    assert 3072 <= H.info.n_coarse <= 4096, H.info.n_coarse
    lap.set_tune({"coarse_prec": prec})
    assert lap.get_tune()["coarse_prec"] == prec
    B = lap.harmonic_rhs(G.seeded_boundary(lap.n_con, k, seed=300 + k))
    Z = O.hsc_apply(O.hsc_build(A.row_ptr, A.col_idx, A.values, coarse_max=4096), B)
    for precision, tol in (("mixed", 2e-5 if prec == 2 else 3e-4), ("fp64", 1e-11)):
        x, info = lap.solve(B, precond="hsc", precision=precision, tol=1e-9, max_iter=1,
                            hsc=H)
        assert info.iters == 1
        X = x.cpu().numpy()
        e = np.linalg.norm(X / np.linalg.norm(X, axis=0) - Z / np.linalg.norm(Z, axis=0),
                           axis=0)
        _say(f"coarse_prec {prec} k {k} {precision}: direction error {e.max():.2e}")
        assert np.all(e <= tol), (precision, e.max())
        x, info = lap.solve(B, precond="hsc", precision=precision, tol=1e-9, hsc=H)
        assert info.status == 0
        R = B - O.spmm(A.row_ptr, A.col_idx, A.values, x.cpu().numpy())
        assert np.all(np.linalg.norm(R, axis=0) <= 1e-9 * np.linalg.norm(B, axis=0))
    with pytest.raises(Exception):
        lap.set_tune({"coarse_prec": 4})


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-11), ("fp64_plain", 1e-11),
                                           ("mixed", 2e-5)])
@pytest.mark.parametrize("k", [3, 100])
def test_pcg_first_step_is_hsc_direction(gpu, precision, tol, k):
    """One PCG step from 0 gives x1 = alpha0 M^-1 b: its direction is the
    oracle hierarchy's apply of b."""
    from tool_godot_curvenet_amd import HSC
    from tool_godot_curvenet_amd import meshgen as G
    lap = _lap("cylinder64")
    A = lap.A_ff()
    B = lap.harmonic_rhs(G.seeded_boundary(lap.n_con, k, seed=60 + k))
    x, info = lap.solve(B, precond="hsc", precision=precision, tol=1e-9, max_iter=1,
                        hsc=HSC(lap))
    assert info.iters == 1
    X = x.cpu().numpy()
    Z = O.hsc_apply(O.hsc_build(A.row_ptr, A.col_idx, A.values), B)
    for c in range(k):
        e = np.linalg.norm(X[:, c] / np.linalg.norm(X[:, c]) - Z[:, c] / np.linalg.norm(Z[:, c]))
        assert e <= tol, (c, e)


# k = 1..4 take the flat streaming kernels (blocks of 768 threads for k = 3,
# 1,024 otherwise) and the narrow level passes; k = 96 the wide ones
@pytest.mark.parametrize("name", ["cylinder64", "icosphere6"])
@pytest.mark.parametrize("precision", ["mixed", "fp64", "fp64_plain"])
@pytest.mark.parametrize("k", [1, 2, 3, 4, 96])
def test_pcg_hsc_converges(gpu, name, precision, k):
    from tool_godot_curvenet_amd import HSC
    from tool_godot_curvenet_amd import meshgen as G
    lap = _lap(name)
    A = lap.A_ff()
    B = lap.harmonic_rhs(G.seeded_boundary(lap.n_con, k, seed=k))
    x, info = lap.solve(B, precond="hsc", precision=precision, tol=1e-9, hsc=HSC(lap))
    assert info.status == 0
    X = x.cpu().numpy()
    R = B - O.spmm(A.row_ptr, A.col_idx, A.values, X)
    rr = np.linalg.norm(R, axis=0) / np.linalg.norm(B, axis=0)
    assert np.all(rr <= 1e-9), rr
    assert info.iters <= 40, info.iters


# A hierarchy of no levels: the coarse block is the whole matrix, and the
# preconditioner its exact inverse.  One FP32 iteration takes the iterated
# residual down to rounding noise, unrelated to the true residual, so the
# column has to restart from the true residual after the replacement rather
# than continue its direction (which diverged to inf before).
@pytest.mark.parametrize("k", [3, 8])
@pytest.mark.parametrize("inner_tol", [0.0316, 1e-4])
def test_pcg_exact_coarse_converges(gpu, k, inner_tol):
    from tool_godot_curvenet_amd import HSC
    from tool_godot_curvenet_amd import meshgen as G
    lap = _lap("cylinder64")
    A = lap.A_ff()
    h = HSC(lap, coarse_max=4096)
    assert h.info.nlevels == 0 and h.info.n_coarse == lap.n_free
    B = lap.harmonic_rhs(G.seeded_boundary(lap.n_con, k, seed=11))
    for precision in ("mixed", "fp64"):
        x, info = lap.solve(B, precond="hsc", precision=precision, tol=1e-9, hsc=h,
                            inner_tol=inner_tol)
        assert info.status == 0, (precision, info.status)
        X = x.cpu().numpy()
        R = B - O.spmm(A.row_ptr, A.col_idx, A.values, X)
        rr = np.linalg.norm(R, axis=0) / np.linalg.norm(B, axis=0)
        assert np.all(rr <= 1e-9), (precision, rr)
        assert info.iters <= 4, (precision, info.iters)


# More than 32 columns that are not a multiple of the vector width (4 FP32,
# 2 FP64) are solved on a copy padded with zero columns: the solution is the
# same bits as solving the padded right-hand side directly, a warm start
# goes through the padding, and the solve converges.
@pytest.mark.parametrize("precision,k,vec", [("mixed", 37, 4), ("mixed", 101, 4),
                                             ("fp64", 37, 2)])
def test_pcg_padded_columns(gpu, precision, k, vec):
    from tool_godot_curvenet_amd import HSC
    from tool_godot_curvenet_amd import meshgen as G
    lap = _lap("icosphere6")
    A = lap.A_ff()
    h = HSC(lap)
    B = lap.harmonic_rhs(G.seeded_boundary(lap.n_con, k, seed=13))
    kp = (k + vec - 1) // vec * vec
    Bp = np.zeros((lap.n_free, kp))
    Bp[:, :k] = B
    x, info = lap.solve(B, precond="hsc", precision=precision, tol=1e-9, hsc=h)
    xp, infop = lap.solve(Bp, precond="hsc", precision=precision, tol=1e-9, hsc=h)
    assert info.status == 0 and infop.status == 0
    X, XP = x.cpu().numpy(), xp.cpu().numpy()
    assert np.array_equal(X, XP[:, :k]) and not XP[:, k:].any()
    assert info.iters == infop.iters
    rr = np.linalg.norm(B - O.spmm(A.row_ptr, A.col_idx, A.values, X), axis=0) / \
        np.linalg.norm(B, axis=0)
    assert np.all(rr <= 1e-9), rr
    # warm start from the solution: no iteration needed
    x2, info2 = lap.solve(B, precond="hsc", precision=precision, tol=1e-9, hsc=h, x0=X)
    assert info2.status == 0 and info2.iters == 0
    assert np.array_equal(x2.cpu().numpy(), X)


def test_hsc_beats_ic0_on_capsules(gpu):
    """The benchmark mesh: far fewer iterations than block IC(0)."""
    from tool_godot_curvenet_amd import HSC, IC0
    from tool_godot_curvenet_amd import meshgen as G
    lap = _lap("capsules100k")
    B = lap.harmonic_rhs(G.seeded_boundary(lap.n_con, 3, seed=5), as_numpy=False)
    _, ih = lap.solve(B, precond="hsc", precision="mixed", tol=1e-9, hsc=HSC(lap))
    _, ii = lap.solve(B, precond="ic0", precision="mixed", tol=1e-9, ic0=IC0(lap))
    assert ih.status == 0 and ii.status == 0
    assert ih.iters * 5 < ii.iters, (ih.iters, ii.iters)


# Applies with many columns run the small levels and the coarse solve in one
# fused launch (one workgroup per column slice): at least 128 slices, i.e.
# k >= 256 in FP64 and k >= 512 in FP32.  On cylinder64 every level is small
# (the whole apply is the fused launch); on icosphere6 the top levels keep
# their own kernels.
@pytest.mark.parametrize("name", ["cylinder64", "icosphere6"])
@pytest.mark.parametrize("k", [256, 301])
def test_hsc_apply_fused_tail(gpu, name, k):
    from tool_godot_curvenet_amd import HSC
    lap = _lap(name)
    A = lap.A_ff()
    H = HSC(lap)
    assert H.tail_level(k, fp32=False) >= 0
    ho = O.hsc_build(A.row_ptr, A.col_idx, A.values)
    r = np.random.default_rng(k).normal(size=(lap.n_free, k))
    zg = H.apply(r)
    zo = O.hsc_apply(ho, r)
    assert np.linalg.norm(zg - zo) <= 1e-11 * np.linalg.norm(zo)


@pytest.mark.parametrize("name", ["cylinder64", "icosphere6"])
@pytest.mark.parametrize("precision,tol", [("mixed", 2e-5), ("fp64", 1e-11)])
def test_pcg_fused_tail_direction_and_convergence(gpu, name, precision, tol):
    from tool_godot_curvenet_amd import HSC
    from tool_godot_curvenet_amd import meshgen as G
    k = 512
    lap = _lap(name)
    A = lap.A_ff()
    H = HSC(lap)
    assert H.tail_level(k, fp32=precision == "mixed") >= 0
    B = lap.harmonic_rhs(G.seeded_boundary(lap.n_con, k, seed=700))
    x, info = lap.solve(B, precond="hsc", precision=precision, tol=1e-9, max_iter=1, hsc=H)
    X = x.cpu().numpy()
    Z = O.hsc_apply(O.hsc_build(A.row_ptr, A.col_idx, A.values), B)
    e = np.linalg.norm(X / np.linalg.norm(X, axis=0) - Z / np.linalg.norm(Z, axis=0), axis=0)
    assert np.all(e <= tol), e.max()
    x, info = lap.solve(B, precond="hsc", precision=precision, tol=1e-9, hsc=H)
    assert info.status == 0
    R = B - O.spmm(A.row_ptr, A.col_idx, A.values, x.cpu().numpy())
    assert np.all(np.linalg.norm(R, axis=0) <= 1e-9 * np.linalg.norm(B, axis=0))
