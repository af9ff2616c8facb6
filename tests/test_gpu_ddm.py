"""GPU Direct Delta Mush kernels against the oracle (pinned in
test_oracle_ddm.py): the skinning matvec (bit-exact in FP64), smoothing and
top-K sparsification (bit-exact), the harvest's right-hand sides (bit-exact
against the harmonic right-hand side of unit vectors) and the bind pipeline's
invariants (partition of unity, unit rows at samples, rigid motions)."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _ddm():
    from tool_godot_curvenet_amd import ddm
    return ddm


def _influences(rng, nv, nt, maxk, bad=True, fixed=False):
    cnt = np.full(nv, maxk) if fixed else rng.integers(0, maxk + 1, size=nv)
    start = np.zeros(nv + 1, np.int32)
    start[1:] = np.cumsum(cnt)
    idx = rng.integers(0, nt, size=start[-1]).astype(np.int32)
    if bad and len(idx) > 2:
        idx[1] = nt + 5                                # reads the zero matrix
        idx[2] = -1
    w = rng.normal(size=start[-1])
    return start, idx, w


# The kernel stages a 256-vertex tile's distinct handles in LDS; tiles with
# more than 8 influences on a vertex, more than 2048 influences in all or
# more distinct handles than the stage holds (512 FP32, 256 FP64) read the
# transforms directly.  Every path keeps the spec's operation order, so each
# must be bit-exact.
@pytest.mark.parametrize("nt,maxk,fixed", [
    (37, 8, False),        # staged in LDS
    (37, 8, True),         # exactly 2048 influences a tile: staged
    (3000, 8, False),      # too many distinct handles a tile: direct loads
    (37, 12, False),       # rows longer than 8: direct loads
    (37, 16, True),        # 4096 influences a tile: direct loads
])
@pytest.mark.parametrize("P", [1, 3, 40])
def test_lbs_f64_bit_exact(gpu, P, nt, maxk, fixed):
    rng = np.random.default_rng(P + nt + maxk)
    nv = 5000
    rest = rng.normal(size=(nv, 3))
    T = rng.normal(size=(P, nt, 16))
    start, idx, w = _influences(rng, nv, nt, maxk, fixed=fixed)
    out = _ddm().lbs_matvec(T, start, idx, w, rest).cpu().numpy()
    for p in range(P):
        assert np.array_equal(out[p], O.lbs_matvec(rest, T[p], start, idx, w)), p


@pytest.mark.parametrize("nt", [64, 3000])
def test_lbs_f32_close(gpu, nt):
    rng = np.random.default_rng(11)
    nv, P = 20000, 4
    rest = rng.normal(size=(nv, 3)).astype(np.float32)
    T = rng.normal(size=(P, nt, 16)).astype(np.float32)
    start, idx, w = _influences(rng, nv, nt, 8, bad=False)
    w = w.astype(np.float32)
    out = _ddm().lbs_matvec(T, start, idx, w, rest, precision="f32").cpu().numpy()
    for p in range(P):
        ref = O.lbs_matvec(rest.astype(np.float64), T[p].astype(np.float64), start, idx,
                           w.astype(np.float64))
        # float rounding of sums of <= 8 products of O(1) terms
        assert np.abs(out[p] - ref).max() <= 1e-5 * max(1.0, np.abs(ref).max())


def test_lbs_kats_on_gpu(gpu):                        # DirectDeltaMush.lean:115-165
    I4 = np.eye(4).reshape(16)
    tr = np.eye(4)
    tr[:3, 3] = (2.0, -3.0, 1.0)
    rest = np.array([[1.5, -2.0, 0.75], [1.0, 1.0, 1.0], [7.0, 8.0, 9.0], [1.0, 1.0, 1.0]])
    T = np.stack([I4, tr.reshape(16), I4, I4, I4])
    start = np.array([0, 1, 2, 6, 6], np.int32)
    idx = np.array([0, 1, 0, 2, 3, 4], np.int32)
    w = np.array([1.0, 1.0, 0.25, 0.25, 0.25, 0.25])
    out = _ddm().lbs_matvec(T, start, idx, w, rest).cpu().numpy()
    assert np.allclose(out, [[1.5, -2.0, 0.75], [3.0, -2.0, 2.0], [7.0, 8.0, 9.0],
                             [0.0, 0.0, 0.0]], atol=1e-12)


def _graph(rng, n, deg):
    nb = [set() for _ in range(n)]
    for _ in range(deg * n // 2):
        a, b = rng.integers(0, n, size=2)
        if a != b:
            nb[a].add(b)
            nb[b].add(a)
    ptr = np.zeros(n + 1, np.int32)
    adj = []
    for v in range(n):
        lst = sorted(nb[v])
        rng.shuffle(lst)
        adj += lst
        ptr[v + 1] = len(adj)
    return ptr, np.array(adj, np.int32)


@pytest.mark.parametrize("nh", [1, 5, 37, 130, 256])
@pytest.mark.parametrize("nu,omega", [(0, 0.5), (1, 0.5), (3, 0.3)])
def test_smooth_bit_exact(gpu, nh, nu, omega):
    rng = np.random.default_rng(nh + nu)
    n = 3000
    ptr, adj = _graph(rng, n, 6)
    adj = adj.copy()
    if len(adj):
        adj[0] = n + 4                                 # out of range: adds 0
    W = rng.random((n, nh))
    got = _ddm().smooth_weights(W, ptr, adj, nu, omega).cpu().numpy()
    assert np.array_equal(got, O.smooth_weights(W, ptr, adj, nu, omega))


@pytest.mark.parametrize("nh,K", [(1, 1), (5, 3), (5, 8), (70, 8), (300, 64), (300, 0),
                                  (64, 64)])
@pytest.mark.parametrize("ties", [False, True])
def test_topk_bit_exact(gpu, nh, K, ties):
    rng = np.random.default_rng(nh * 7 + K)
    n = 2000
    if ties:
        W = rng.integers(-3, 4, size=(n, nh)).astype(np.float64) / 4.0
    else:
        W = rng.normal(size=(n, nh))
    idx, w = _ddm().sparsify_topk(W, K)
    ri, rw = O.sparsify_topk(W, K)
    assert np.array_equal(idx.cpu().numpy(), ri)
    assert np.array_equal(w.cpu().numpy(), rw)


def test_topk_nan_and_limits(gpu):
    from tool_godot_curvenet_amd._lib import CurvenetError
    W = np.array([[np.nan, 1.0, 2.0, 0.5], [1.0, np.nan, 3.0, -4.0]])
    idx, w = _ddm().sparsify_topk(W, 3)
    ri, rw = O.sparsify_topk(W, 3)
    assert np.array_equal(idx.cpu().numpy(), ri)
    assert np.array_equal(w.cpu().numpy(), rw, equal_nan=True)
    with pytest.raises(CurvenetError):
        _ddm().sparsify_topk(np.zeros((2, 100)), 65)


def _lap(name):
    from tool_godot_curvenet_amd import CutLaplacian, meshgen
    m = meshgen.config_mesh(name)
    return m, CutLaplacian.from_mesh(m)


def test_bind_rhs_equals_unit_harmonic_rhs(gpu):
    import torch
    from tool_godot_curvenet_amd import _lib as L
    from tool_godot_curvenet_amd.curvenet import _p, _stream
    _, lap = _lap("cylinder64")
    h0, nb = 5, 40
    out = torch.empty((lap.n_free, nb), dtype=torch.float64, device="cuda")
    L.check(L.load().cn_bind_rhs(lap.handle, h0, nb, _p(out), _stream()), "bind_rhs")
    E = np.zeros((lap.n_con, nb))
    E[h0 + np.arange(nb), np.arange(nb)] = 1.0
    got = out.cpu().numpy()
    assert np.array_equal(got, lap.harmonic_rhs(E))
    h = O.he_from_triangles(lap.pos.shape[0], lap.tris)
    free_of, con_of = lap.maps()
    assert np.array_equal(got, O.harmonic_rhs_csr(h, free_of, con_of, lap.pos, E))
    with pytest.raises(L.CurvenetError):
        L.check(L.load().cn_bind_rhs(lap.handle, lap.n_con - 3, 4, _p(out), _stream()), "x")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_bind_harvest_invariants(gpu, precision):
    _, lap = _lap("cylinder64")
    W, infos = _ddm().bind_harvest(lap, block=48, precision=precision, tol=1e-10)
    assert all(i.status == 0 for i in infos)
    Wn = W.cpu().numpy()
    free_of, con_of = lap.maps()
    # sample vertices carry their own handle with weight 1
    s = con_of >= 0
    assert np.array_equal(Wn[s], np.eye(lap.n_con)[con_of[s]])
    # harmonic: A_ff X + A_fc (I) = 0 on the free rows
    A, Afc = lap.A_ff(), lap.A_fc()
    sel = free_of >= 0
    fv = np.empty(lap.n_free, np.int64)
    fv[free_of[sel]] = np.nonzero(sel)[0]
    X = Wn[fv]
    R = O.spmm(A.row_ptr, A.col_idx, A.values, X) + O.spmm(
        Afc.row_ptr, Afc.col_idx, Afc.values, np.eye(lap.n_con))
    B = lap.harmonic_rhs(np.eye(lap.n_con))
    assert np.all(np.linalg.norm(R, axis=0) <= 1e-9 * np.linalg.norm(B, axis=0))
    # constants are harmonic: rows sum to 1 (partition of unity)
    assert _ddm().partition_of_unity(W, 1e-7)


def test_bind_pipeline_rigid_motions(gpu):
    """Harvest -> smooth -> top-K -> LBS: identity transforms give the rest
    pose, one common rigid motion moves every vertex by it."""
    import torch
    m, lap = _lap("cylinder64")
    start, idx, w, infos = _ddm().bind(lap, lap.tris, nu=2, omega=0.5, K=8,
                                       precision="fp64", tol=1e-11)
    assert all(i.status == 0 for i in infos)
    H = lap.n_con
    assert idx.numel() == 8 * lap.info.nv
    # every kept row sums to 1 (the harvest is a partition of unity)
    ws = w.reshape(-1, 8).sum(dim=1)
    assert torch.allclose(ws, torch.ones_like(ws), atol=1e-9)
    rest = lap.pos
    I = np.tile(np.eye(4).reshape(16), (H, 1))
    out = _ddm().lbs_matvec(I, start, idx, w, rest).cpu().numpy()
    assert np.allclose(out, rest, atol=1e-9)
    c, s_ = np.cos(0.4), np.sin(0.4)
    M = np.eye(4)
    M[:3, :3] = [[c, 0, s_], [0, 1, 0], [-s_, 0, c]]
    M[:3, 3] = (0.3, -1.0, 2.0)
    out = _ddm().lbs_matvec(np.tile(M.reshape(16), (H, 1)), start, idx, w, rest).cpu().numpy()
    assert np.allclose(out, rest @ M[:3, :3].T + M[:3, 3], atol=1e-8)
    # the top-K output is the oracle's on the same smoothed matrix
    W, _ = _ddm().bind_harvest(lap, precision="fp64", tol=1e-11)
    ap, aj = _ddm().vertex_adjacency(lap.info.nv, lap.tris)
    Ws = O.smooth_weights(W.cpu().numpy(), ap, aj, 2, 0.5)
    ri, rw = O.sparsify_topk(Ws, 8)
    assert np.array_equal(_ddm().sparsify_topk(Ws, 8)[0].cpu().numpy(), ri)
