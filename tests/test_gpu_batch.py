"""Batches of independent systems (cn_system_batch): one block-diagonal
system whose (member, column) pairs follow their own CG recurrences, so each
member's rows reproduce that member's own solve."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _members():
    from tool_godot_curvenet_amd import CutLaplacian
    from tool_godot_curvenet_amd import meshgen as G
    meshes = [G.open_cylinder(24, 16, seed=1),
              G.tilted_loops(G.geodesic_sphere(12, jitter_frac=0.05, seed=2), 4, 30.0, seed=3),
              G.capsule_union(2, 3, 3000, seed=4),
              G.open_cylinder(40, 30, seed=5)]
    return [CutLaplacian.from_mesh(m) for m in meshes]


@pytest.fixture(scope="module")
def batch():
    from tool_godot_curvenet_amd import BatchSystem
    mem = _members()
    return mem, BatchSystem(mem)


def _rhs(mem, k, seed):
    from tool_godot_curvenet_amd import meshgen as G
    return [m.harmonic_rhs(G.seeded_boundary(m.n_con, k, seed=seed + i)) for i, m in enumerate(mem)]


def test_batch_is_block_diagonal(gpu, batch):
    mem, B = batch
    A = B.A_ff()
    assert B.nseg == len(mem)
    assert list(np.diff(B.seg_ptr)) == [m.n_free for m in mem]
    r0, z0 = 0, 0
    for i, m in enumerate(mem):
        M = m.A_ff()
        a, b = B.seg_ptr[i], B.seg_ptr[i + 1]
        assert np.array_equal(A.row_ptr[a:b + 1] - A.row_ptr[a], M.row_ptr)
        seg = slice(A.row_ptr[a], A.row_ptr[b])
        assert np.array_equal(A.col_idx[seg] - r0, M.col_idx)
        assert np.array_equal(A.values[seg], M.values)
        r0 += m.n_free
        z0 += M.nnz
    assert A.nnz == z0


@pytest.mark.parametrize("precond", ["jacobi", "sgs"])
@pytest.mark.parametrize("k", [1, 2, 3, 4, 40])
def test_two_step_iterates_match_members(gpu, batch, precond, k):
    """After two FP64 iterations every member's rows equal its own solve's
    iterate (alpha, beta per member and column: shared scalars would give
    other iterates); only the dots' summation order differs."""
    from tool_godot_curvenet_amd import SGS
    mem, B = batch
    rhs = _rhs(mem, k, 10)
    kw = {"sgs": SGS(B)} if precond == "sgs" else {}
    xb, ib = B.solve(B.stack(rhs), precond=precond, precision="fp64", tol=1e-12, max_iter=2, **kw)
    assert ib.iters == 2
    for m, b, x in zip(mem, rhs, B.split(xb.cpu().numpy())):
        kwm = {"sgs": SGS(m)} if precond == "sgs" else {}
        xm, im = m.solve(b, precond=precond, precision="fp64", tol=1e-12, max_iter=2, **kwm)
        xm = xm.cpu().numpy()
        assert np.abs(x - xm).max() <= 1e-12 * np.abs(xm).max()


@pytest.mark.parametrize("precond", ["hsc", "sgs", "jacobi"])
@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_batch_solves_every_member(gpu, batch, precond, precision):
    from tool_godot_curvenet_amd import HSC, SGS
    mem, B = batch
    rhs = _rhs(mem, 3, 20)
    kw = {"hsc": HSC(B)} if precond == "hsc" else ({"sgs": SGS(B)} if precond == "sgs" else {})
    x, info = B.solve(B.stack(rhs), precond=precond, precision=precision, tol=1e-9, **kw)
    assert info.status == 0
    for m, b, xm in zip(mem, rhs, B.split(x.cpu().numpy())):
        A = m.A_ff()
        R = b - O.spmm(A.row_ptr, A.col_idx, A.values, xm)
        rr = np.linalg.norm(R, axis=0) / np.linalg.norm(b, axis=0)
        assert np.all(rr <= 1e-9), rr


def test_batch_rejects_ic0_and_plain(gpu, batch):
    from tool_godot_curvenet_amd._lib import CurvenetError
    mem, B = batch
    b = B.stack(_rhs(mem, 3, 30))
    with pytest.raises(CurvenetError):
        B.solve(b, precond="jacobi", precision="fp64_plain")
    with pytest.raises((CurvenetError, Exception)):
        B.solve(b, precond="ic0", precision="mixed")


def test_batch_harmonic_rhs_one_launch(gpu):
    """cn_harmonic_rhs on a batch builds every member's right-hand side from
    the stacked constrained values: the same bits as the member calls."""
    from tool_godot_curvenet_amd import BatchSystem, CutLaplacian, meshgen
    laps = [CutLaplacian.from_mesh(meshgen.capsule_union(c, c + 2, t, seed=31 + c))
            for c, t in ((1, 2500), (2, 7000), (1, 4000))]
    batch = BatchSystem(laps)
    rng = np.random.default_rng(2)
    for k in (1, 3, 8):
        Xc = [rng.uniform(-0.2, 0.2, (l.n_con, k)) for l in laps]
        got = batch.harmonic_rhs(Xc)
        want = batch.stack([l.harmonic_rhs(x) for l, x in zip(laps, Xc)])
        assert np.array_equal(got, want), k
