"""GPU deformation solve (DeformSolve.solveDeformation) and its curve
inputs against the oracle (pinned in test_oracle_deform.py)."""
import math

import numpy as np
import pytest

import oracle as O
from kat_data import EQ_TRI

pytestmark = pytest.mark.gpu

R90 = np.array([[0.0, -1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 1.0]])


def _lap(name):
    from tool_godot_curvenet_amd import CutLaplacian, meshgen
    return CutLaplacian.from_mesh(meshgen.config_mesh(name))


@pytest.mark.parametrize("name", ["cylinder64", "icosphere6"])
def test_deform_rhs_bit_exact(gpu, name):
    import torch
    from tool_godot_curvenet_amd import _lib as L
    from tool_godot_curvenet_amd.curvenet import _p, _stream
    lap = _lap(name)
    h = O.he_from_triangles(lap.pos.shape[0], lap.tris)
    free_of, con_of = lap.maps()
    rng = np.random.default_rng(3)
    P = 2
    Ff = rng.normal(size=(lap.n_free, 9 * P))
    Fc = rng.normal(size=(lap.n_con, 9 * P))
    Xc = rng.normal(size=(lap.n_con, 3 * P))
    t = [torch.as_tensor(a, device="cuda") for a in (Ff, Fc, Xc)]
    out = torch.empty((lap.n_free, 3 * P), dtype=torch.float64, device="cuda")
    L.check(L.load().cn_deform_rhs(lap.handle, P, _p(t[0]), _p(t[1]), _p(t[2]), _p(out),
                                   _stream()), "deform_rhs")
    want = O.deform_rhs_csr(h, free_of, con_of, lap.pos, Ff, Fc, Xc)
    assert np.array_equal(out.cpu().numpy(), want)


def _tri_with_sample():
    from tool_godot_curvenet_amd import CutLaplacian
    return CutLaplacian(EQ_TRI, np.array([[0, 1, 2]], np.int32), np.array([1, 0, 0], np.uint8))


@pytest.mark.parametrize("Fc,Xc,v1,v2", [
    (np.eye(3), [5.0, 0.0, 0.0], [6.0, 0.0, 0.0], [5.5, 0.0, 0.8660254037844386]),
    (np.eye(3), [0.0, 2.0, 0.0], [1.0, 2.0, 0.0], None),
    (R90, [0.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.5, 0.8660254037844386]),
    (2 * np.eye(3), [0.0, 0.0, 0.0], [2.0, 0.0, 0.0], [1.0, 0.0, 1.7320508075688772]),
])
def test_deform_solve_kats(gpu, Fc, Xc, v1, v2):       # DeformSolve.lean:164-253
    from tool_godot_curvenet_amd import deform_solve
    lap = _tri_with_sample()
    _, xa, i1, i2 = deform_solve(lap, np.reshape(Fc, (1, 9)), np.reshape(Xc, (1, 3)),
                                 precond="jacobi", precision="fp64", tol=1e-13)
    X = xa.cpu().numpy()
    assert np.allclose(X[1], v1, atol=1e-9)
    if v2 is not None:
        assert np.allclose(X[2], v2, atol=1e-9)
    assert np.allclose(X[0], Xc, atol=0)                 # the sample, overlaid


def test_bezier_driven_translation_end_to_end(gpu):     # EndToEndExample.lean:62-211
    from tool_godot_curvenet_amd import bezier3, deform_solve, segment_gradients
    rest = np.array([[-1.5, 0, 0], [-0.5, 0, 0], [0.5, 0, 0], [1.5, 0, 0]])
    ctrl = np.stack([rest, rest + [5.0, 0, 0], rest @ R90.T])
    pts = bezier3(ctrl, [0.5, 0.6])                     # 3 curves x 2 parameters
    for c in range(3):
        for k, t in enumerate((0.5, 0.6)):
            assert np.array_equal(pts[c, k], O.bezier3(ctrl[c], t))
    F = segment_gradients(pts[[0, 0], 0], pts[[0, 0], 1], pts[[1, 2], 0], pts[[1, 2], 1])
    lap = _tri_with_sample()
    _, xa, _, _ = deform_solve(lap, F[0].reshape(1, 9), pts[1, 0].reshape(1, 3),
                               precond="jacobi", precision="fp64", tol=1e-13)
    X = xa.cpu().numpy()
    assert np.allclose(X[1], [6.0, 0, 0], atol=1e-9)
    assert np.allclose(X[2], [5.5, 0, 0.8660254037844386], atol=1e-9)
    _, xr, _, _ = deform_solve(lap, F[1].reshape(1, 9), pts[2, 0].reshape(1, 3),
                               precond="jacobi", precision="fp64", tol=1e-13)
    X = xr.cpu().numpy()
    assert np.allclose(X[1], [0.0, 1.0, 0], atol=1e-9)
    assert np.allclose(X[2], [0.0, 0.5, 0.8660254037844386], atol=1e-9)


def test_curve_kernels_bit_exact(gpu):
    from tool_godot_curvenet_amd import deformation_gradients, segment_gradients
    rng = np.random.default_rng(7)
    n = 500
    a, b, c, d = rng.normal(size=(4, n, 3))
    F = segment_gradients(a, b, c, d)
    for s in range(0, n, 7):
        assert np.array_equal(F[s], O.isolated_segment_gradient(a[s], b[s], c[s], d[s]))
    rest, posed = rng.normal(size=(2, n, 3, 3))
    G = deformation_gradients(rest, posed)
    for s in range(0, n, 7):
        assert np.array_equal(G[s], O.deformation_gradient(rest[s], posed[s]))


@pytest.mark.parametrize("name", ["cylinder64", "sphere45"])
def test_deform_solve_reproduces_rigid_motions(gpu, name):
    """F_c = R and X_c = R x_c + t on every constrained vertex: every vertex
    moves rigidly (three poses at once, HSC-preconditioned)."""
    from tool_godot_curvenet_amd import HSC, deform_solve
    lap = _lap(name)
    free_of, con_of = lap.maps()
    rng = np.random.default_rng(11)
    Rs, ts = [], []
    for _ in range(3):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        w, x, y, z = q
        Rs.append(np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                            [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                            [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]]))
        ts.append(rng.normal(size=3))
    cidx = np.flatnonzero(con_of >= 0)
    Fc = np.zeros((lap.n_con, 27))
    Xc = np.zeros((lap.n_con, 9))
    for p in range(3):
        Fc[:, 9 * p:9 * p + 9] = Rs[p].reshape(1, 9)
        Xc[con_of[cidx], 3 * p:3 * p + 3] = lap.pos[cidx] @ Rs[p].T + ts[p]
    _, xa, i1, i2 = deform_solve(lap, Fc, Xc, precond="hsc", precision="fp64", tol=1e-12,
                                 hsc=HSC(lap))
    assert i1.status == 0 and i2.status == 0
    X = xa.cpu().numpy()
    for p in range(3):
        want = lap.pos @ Rs[p].T + ts[p]
        assert np.abs(X[:, 3 * p:3 * p + 3] - want).max() < 1e-8 * np.abs(want).max()
