"""GPU colour-parallel symmetric Gauss-Seidel, the SGS-preconditioned PCG
and the coarse-vertex restriction, against the oracle (pinned in
test_oracle_sgs.py)."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _lap(name):
    from tool_godot_curvenet_amd import CutLaplacian, meshgen
    return CutLaplacian.from_mesh(meshgen.config_mesh(name))


@pytest.fixture(scope="module", params=["cylinder64", "icosphere6"])
def sys_(request):
    from tool_godot_curvenet_amd import SGS
    lap = _lap(request.param)
    return lap, lap.A_ff(), SGS(lap)


def test_colouring_matches_oracle(gpu, sys_):
    lap, A, G = sys_
    color, nc = O.greedy_color(A.row_ptr, A.col_idx)
    cptr, rows = O.color_classes(color, nc)
    assert G.ncolors == nc
    assert np.array_equal(G.color, color)
    assert np.array_equal(G.color_ptr, cptr) and np.array_equal(G.color_rows, rows)


@pytest.mark.parametrize("k,sweeps", [(1, 1), (3, 2), (37, 1)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_sweep_bit_exact(gpu, sys_, k, sweeps, dtype):
    """The row update has no fused multiply-add and sums in storage order:
    the same bits as the oracle's loop, in FP64 and in FP32."""
    lap, A, G = sys_
    rng = np.random.default_rng(k + sweeps)
    b = rng.normal(size=(lap.n_free, k))
    x0 = rng.normal(size=(lap.n_free, k))
    xg = G.sweep(b, x0, sweeps, dtype=dtype)
    xo = O.sgs_sweep(A.row_ptr, A.col_idx, A.values, b, x0, sweeps, dtype=dtype)
    assert xg.dtype == dtype
    assert np.array_equal(xg, xo)


def test_apply_bit_exact(gpu, sys_):
    lap, A, G = sys_
    r = np.random.default_rng(5).normal(size=(lap.n_free, 4))
    assert np.array_equal(G.apply(r), O.sgs_apply(A.row_ptr, A.col_idx, A.values, r))


def test_zero_diagonal_row_untouched(gpu):
    from tool_godot_curvenet_amd import SGS, CsrMatrix, CsrSystem
    ptr = np.array([0, 2, 5, 7], np.int32)
    col = np.array([0, 1, 0, 1, 2, 1, 2], np.int32)
    val = np.array([2.0, 1.0, 1.0, 0.0, 1.0, 1.0, 2.0])
    S = CsrSystem(CsrMatrix(3, 3, ptr, col, val))
    x0 = np.array([0.5, 7.0, -0.5])
    x = SGS(S).sweep(np.ones(3), x0, 3)
    assert np.array_equal(x, O.sgs_sweep(ptr, col, val, np.ones(3), x0, 3))
    assert x[1] == 7.0


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-11), ("fp64_plain", 1e-11),
                                           ("mixed", 2e-5)])
@pytest.mark.parametrize("k", [3, 100])
def test_pcg_first_step_is_sgs_direction(gpu, precision, tol, k):
    """One PCG step from 0: x1 = alpha0 M^-1 b, M^-1 b from the oracle."""
    from tool_godot_curvenet_amd import SGS
    from tool_godot_curvenet_amd import meshgen as G
    lap = _lap("cylinder64")
    A = lap.A_ff()
    B = lap.harmonic_rhs(G.seeded_boundary(lap.n_con, k, seed=70 + k))
    x, info = lap.solve(B, precond="sgs", precision=precision, tol=1e-9, max_iter=1,
                        sgs=SGS(lap))
    assert info.iters == 1
    X = x.cpu().numpy()
    Z = O.sgs_apply(A.row_ptr, A.col_idx, A.values, B)
    for c in range(k):
        e = np.linalg.norm(X[:, c] / np.linalg.norm(X[:, c]) - Z[:, c] / np.linalg.norm(Z[:, c]))
        assert e <= tol, (c, e)


@pytest.mark.parametrize("name", ["cylinder64", "icosphere6"])
@pytest.mark.parametrize("precision", ["mixed", "fp64", "fp64_plain"])
@pytest.mark.parametrize("k", [3, 96])
def test_pcg_sgs_converges(gpu, name, precision, k):
    from tool_godot_curvenet_amd import SGS
    from tool_godot_curvenet_amd import meshgen as G
    lap = _lap(name)
    A = lap.A_ff()
    B = lap.harmonic_rhs(G.seeded_boundary(lap.n_con, k, seed=k + 1))
    x, info = lap.solve(B, precond="sgs", precision=precision, tol=1e-9, sgs=SGS(lap))
    assert info.status == 0
    R = B - O.spmm(A.row_ptr, A.col_idx, A.values, x.cpu().numpy())
    rr = np.linalg.norm(R, axis=0) / np.linalg.norm(B, axis=0)
    assert np.all(rr <= 1e-9), rr
    _, ij = lap.solve(B[:, :3], precond="jacobi", precision="fp64", tol=1e-9)
    assert info.iters < ij.iters          # SGS beats Jacobi


@pytest.mark.parametrize("k", [1, 5, 300])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_restriction_bit_exact(gpu, k, dtype):
    from tool_godot_curvenet_amd import Restriction
    rng = np.random.default_rng(k)
    n_fine, n_coarse = 5000, 700
    parent = rng.integers(-1, n_coarse + 20, size=n_fine).astype(np.int32)
    fine = rng.normal(size=(n_fine, k))
    got = Restriction(parent, n_coarse)(fine, dtype=dtype)
    want = O.restrict_sum(parent, n_coarse, fine, dtype=dtype)
    assert got.dtype == dtype and np.array_equal(got, want)


def test_restriction_kat(gpu):
    from tool_godot_curvenet_amd import Restriction
    parent = np.array([0, 1, 0, 2, 1, -1, 3], np.int32)
    got = Restriction(parent, 4)(np.array([1.0, 2.0, 3.0, 4.0, 5.0, 6.0, 7.0]))
    assert np.array_equal(got, [4.0, 7.0, 4.0, 7.0])
