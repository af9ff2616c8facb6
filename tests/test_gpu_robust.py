"""GPU robust (Sharp-Crane) Laplacian assembly against the oracle (pinned
in test_oracle_robust.py): bit-exact CSR and right-hand sides, finite on a
mesh with a collinear triangle where the embedded cot is not."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _check_bit_exact(lap, delta):
    h = O.he_from_triangles(lap.pos.shape[0], lap.tris)
    with O.laplacian("robust", delta):
        o = O.cut_laplacian_csr(h, lap.constrained, lap.pos)
    A = lap.A_ff()
    for got, want in zip((A.row_ptr, A.col_idx, A.values), o["A_ff"]):
        assert np.array_equal(got, want)
    B = lap.A_fc()
    for got, want in zip((B.row_ptr, B.col_idx, B.values), o["A_fc"]):
        assert np.array_equal(got, want)
    return h, o


@pytest.mark.parametrize("name", ["cylinder64", "icosphere6"])
@pytest.mark.parametrize("delta", [0.0, 0.01])
def test_robust_assembly_bit_exact(gpu, name, delta):
    from tool_godot_curvenet_amd import CutLaplacian, meshgen as G
    lap = CutLaplacian.from_mesh(G.config_mesh(name), laplacian="robust", delta=delta)
    h, o = _check_bit_exact(lap, delta)
    Xc = G.seeded_boundary(lap.n_con, 3)
    with O.laplacian("robust", delta):
        want = O.harmonic_rhs_csr(h, o["free_of"], o["con_of"], lap.pos, Xc)
    assert np.array_equal(lap.harmonic_rhs(Xc), want)


def _sliver_cylinder():
    """A 24 x 12 cylinder with one vertex moved onto a neighbour (a
    triangle with a zero-length edge: its embedded cot is 0/0)."""
    from tool_godot_curvenet_amd import meshgen as G
    m = G.open_cylinder(24, 12, seed=11)
    pos = m.pos.copy()
    t = m.tris[24 * 3 * 2 + 3]                   # a triangle below the seam ring
    pos[t[0]] = pos[t[1]]
    return G.Mesh(pos, m.tris, m.cut_edges, "sliver", dict(m.meta))


def test_robust_is_finite_where_cot_is_not(gpu):
    from tool_godot_curvenet_amd import CutLaplacian
    m = _sliver_cylinder()
    cot = CutLaplacian.from_mesh(m)
    assert not np.all(np.isfinite(cot.A_ff().values))
    rob = CutLaplacian.from_mesh(m, laplacian="robust", delta=0.01)
    assert np.all(np.isfinite(rob.A_ff().values))
    _check_bit_exact(rob, 0.01)


def test_robust_system_solves_and_survives_new_positions(gpu):
    from tool_godot_curvenet_amd import CutLaplacian, meshgen as G
    m = _sliver_cylinder()
    lap = CutLaplacian.from_mesh(m, laplacian="robust", delta=0.01)
    A = lap.A_ff()
    B = lap.harmonic_rhs(G.seeded_boundary(lap.n_con, 3, seed=2))
    x, info = lap.solve(B, precond="jacobi", precision="fp64", tol=1e-9)
    assert info.status == 0
    R = B - O.spmm(A.row_ptr, A.col_idx, A.values, x.cpu().numpy())
    assert np.all(np.linalg.norm(R, axis=0) <= 1e-9 * np.linalg.norm(B, axis=0))
    lap.set_positions(lap.pos * np.array([1.0, 1.3, 0.8]))
    _check_bit_exact(lap, 0.01)
