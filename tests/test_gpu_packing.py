"""Moderate column counts (5 <= k <= 200), where the wide row kernels pack
12, 16 or 32 lanes per row: a few PCG iterations against the same iterations
in numpy over the oracle's SpMM and hierarchy, and converged solves.

The lanes per row follow from k: k = 5, 12, 48, 144 in FP32 and k = 5, 12 in
FP64 pack 12 lanes per row (5 rows a wavefront, 4 lanes idle); k = 27 and 190
pack 32 (190 in three column tiles), and so does k = 48 in FP64; k = 144 and
200 in FP64 pack 16; k = 200 in FP32 keeps 64.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

KS = [5, 12, 27, 48, 144, 190, 200]


def _lap(name):
    from tool_godot_curvenet_amd import CutLaplacian, meshgen
    return CutLaplacian.from_mesh(meshgen.config_mesh(name))


def _pcg_steps(A, B, apply_m, iters):
    """`iters` steps of preconditioned CG from 0 (SparseLinAlg.lean:104-163
    with M^-1 applied to the residual), every column at once."""
    X = np.zeros_like(B)
    R = B.copy()
    Z = apply_m(R)
    P = Z.copy()
    rz = np.sum(R * Z, axis=0)
    for _ in range(iters):
        Q = O.spmm(A.row_ptr, A.col_idx, A.values, P)
        al = rz / np.sum(P * Q, axis=0)
        X += al * P
        R -= al * Q
        Z = apply_m(R)
        rz1 = np.sum(R * Z, axis=0)
        P = Z + (rz1 / rz) * P
        rz = rz1
    return X


@pytest.mark.parametrize("precond", ["hsc", "jacobi"])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-9), ("mixed", 2e-4)])
@pytest.mark.parametrize("k", KS)
def test_pcg_steps_match_numpy(gpu, precond, precision, tol, k):
    from tool_godot_curvenet_amd import HSC
    from tool_godot_curvenet_amd import meshgen as G
    lap = _lap("cylinder64")
    A = lap.A_ff()
    B = lap.harmonic_rhs(G.seeded_boundary(lap.n_con, k, seed=300 + k))
    if precond == "hsc":
        h = O.hsc_build(A.row_ptr, A.col_idx, A.values)
        apply_m = lambda R: O.hsc_apply(h, R)  # noqa: E731
        extra = {"hsc": HSC(lap)}
    else:
        d = O.csr_diag(A.row_ptr, A.col_idx, A.values)
        dinv = np.where(d != 0.0, 1.0 / np.where(d != 0.0, d, 1.0), 0.0)
        apply_m = lambda R: R * dinv[:, None]  # noqa: E731
        extra = {}
    x, info = lap.solve(B, precond=precond, precision=precision, tol=1e-14, max_iter=3,
                        **extra)
    assert info.iters == 3, info.iters
    X = x.cpu().numpy()
    Xo = _pcg_steps(A, B, apply_m, 3)
    err = np.linalg.norm(X - Xo, axis=0) / np.linalg.norm(Xo, axis=0)
    assert np.all(err <= tol), (err.max(), int(err.argmax()))


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
@pytest.mark.parametrize("k", [9, 27, 48, 144])
def test_pcg_hsc_converges_moderate_k(gpu, precision, k):
    from tool_godot_curvenet_amd import HSC
    from tool_godot_curvenet_amd import meshgen as G
    lap = _lap("icosphere6")
    A = lap.A_ff()
    B = lap.harmonic_rhs(G.seeded_boundary(lap.n_con, k, seed=400 + k))
    x, info = lap.solve(B, precond="hsc", precision=precision, tol=1e-9, hsc=HSC(lap))
    assert info.status == 0
    X = x.cpu().numpy()
    R = B - O.spmm(A.row_ptr, A.col_idx, A.values, X)
    rr = np.linalg.norm(R, axis=0) / np.linalg.norm(B, axis=0)
    assert np.all(rr <= 1e-9), rr.max()


@pytest.mark.parametrize("k", [5, 12, 27, 48, 144])
def test_hsc_apply_moderate_k(gpu, k):
    from tool_godot_curvenet_amd import HSC
    lap = _lap("icosphere6")
    A = lap.A_ff()
    H = HSC(lap)
    ho = O.hsc_build(A.row_ptr, A.col_idx, A.values)
    r = np.random.default_rng(k).normal(size=(lap.n_free, k))
    zg = H.apply(r)
    zo = O.hsc_apply(ho, r)
    assert np.linalg.norm(zg - zo) <= 1e-11 * np.linalg.norm(zo)


def test_warm_start_moderate_k(gpu):
    """A warm start runs the FP64 residual kernel on a nonzero x (k = 27:
    32 lanes per row)."""
    from tool_godot_curvenet_amd import HSC
    from tool_godot_curvenet_amd import meshgen as G
    lap = _lap("cylinder64")
    A = lap.A_ff()
    k = 27
    B = lap.harmonic_rhs(G.seeded_boundary(lap.n_con, k, seed=77))
    H = HSC(lap)
    x1, _ = lap.solve(B, precond="hsc", precision="fp64", tol=1e-4, hsc=H)
    x, info = lap.solve(B, precond="hsc", precision="mixed", tol=1e-10, hsc=H, x0=x1)
    assert info.status == 0
    X = x.cpu().numpy()
    R = B - O.spmm(A.row_ptr, A.col_idx, A.values, X)
    rr = np.linalg.norm(R, axis=0) / np.linalg.norm(B, axis=0)
    assert np.all(rr <= 1e-10), rr.max()


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("mixed", 1e-5)])
@pytest.mark.parametrize("name,k", [("capsules100k", 768), ("icosphere6", 96), ("cylinder64", 8)])
def test_spmm_forms_agree(gpu, precision, tol, name, k):
    """The LDS-staged SpMM and the row-table SpMM compute the same products
    in the same order (only the p.q partial sums are grouped differently):
    two PCG iterations agree to rounding."""
    from tool_godot_curvenet_amd import CutLaplacian, HSC, meshgen
    lap = CutLaplacian.from_mesh(meshgen.config_mesh(name))
    B = lap.harmonic_rhs(meshgen.seeded_boundary(lap.n_con, k, seed=k))
    h = HSC(lap)
    import ctypes
    from tool_godot_curvenet_amd import _lib
    lib = _lib.load()
    xs = {}
    for form in ("row", "lds"):
        lap.set_tune({"spmm_form": 1 if form == "row" else 2})
        lib.cn_profile_enable(1)
        x, info = lap.solve(B, precond="hsc", precision=precision, tol=1e-9, max_iter=2, hsc=h)
        ms = ctypes.c_double(0.0)
        # the LDS-staged product alone, or the fused pair of them
        used = lib.cn_profile_query(b"spmm_lds", ctypes.byref(ms)) + \
            lib.cn_profile_query(b"spmm_dir", ctypes.byref(ms))
        lib.cn_profile_enable(0)
        assert (used > 0) == (form == "lds"), (form, used)
        assert info.iters == 2
        xs[form] = x.cpu().numpy()
    d = np.abs(xs["row"] - xs["lds"]).max() / np.abs(xs["row"]).max()
    assert d <= tol, d


@pytest.mark.parametrize("name,k,precond,precision", [
    ("capsules100k", 768, "hsc", "mixed"),
    ("icosphere6", 48, "hsc", "fp64"),
    ("icosphere6", 3, "hsc", "mixed"),
    ("cylinder64", 3, "ic0", "mixed"),
    ("cylinder64", 5, "jacobi", "fp64"),
])
def test_graph_replay_bit_identical(gpu, name, k, precond, precision):
    """The inner iterations replayed from a captured hipGraph run the same
    kernels in the same order as direct launches: the same bits, also on
    a second solve that replays the cached graph."""
    from tool_godot_curvenet_amd import HSC, IC0, CutLaplacian, meshgen
    lap = CutLaplacian.from_mesh(meshgen.config_mesh(name))
    B = lap.harmonic_rhs(meshgen.seeded_boundary(lap.n_con, k, seed=3 * k))
    kw = {"hsc": HSC(lap)} if precond == "hsc" else ({"ic0": IC0(lap)} if precond == "ic0"
                                                     else {})
    x0, i0 = lap.solve(B, precond=precond, precision=precision, tol=1e-9, graph=False, **kw)
    for rep in range(2):
        x1, i1 = lap.solve(B, precond=precond, precision=precision, tol=1e-9, graph=True, **kw)
        assert i1.status == 0 and i1.iters == i0.iters
        assert np.array_equal(x0.cpu().numpy(), x1.cpu().numpy()), rep
    A = lap.A_ff()
    R = B - O.spmm(A.row_ptr, A.col_idx, A.values, x1.cpu().numpy())
    assert np.all(np.linalg.norm(R, axis=0) <= 1e-9 * np.linalg.norm(B, axis=0))


def test_graph_replay_split_coarse_product(gpu):
    """The benched capsule hierarchy (theta 0.002, 3,834 coarse rows) with the
    split-bfloat16 coarse product (coarse_prec 3) at k = 768: graph replay
    gives the direct launches' bits, and the solve reaches 1e-9."""
    from tool_godot_curvenet_amd import HSC, CutLaplacian, meshgen
    lap = CutLaplacian.from_mesh(meshgen.config_mesh("capsules100k"))
    h = HSC(lap, theta=0.002, coarse_max=4096)
    assert h.info.n_coarse >= 3072
    lap.set_tune({"spmm_form": 2, "coarse_prec": 3})
    B = lap.harmonic_rhs(meshgen.seeded_boundary(lap.n_con, 768, seed=77))
    x0, i0 = lap.solve(B, precond="hsc", precision="mixed", tol=1e-9, graph=False, hsc=h)
    for rep in range(2):
        x1, i1 = lap.solve(B, precond="hsc", precision="mixed", tol=1e-9, graph=True, hsc=h)
        assert i1.status == 0 and i1.iters == i0.iters
        assert np.array_equal(x0.cpu().numpy(), x1.cpu().numpy()), rep
    A = lap.A_ff()
    R = B - O.spmm(A.row_ptr, A.col_idx, A.values, x1.cpu().numpy())
    assert np.all(np.linalg.norm(R, axis=0) <= 1e-9 * np.linalg.norm(B, axis=0))


def test_graph_replay_batch(gpu):
    """A batch of systems (per-member scalars): replayed equals direct."""
    from tool_godot_curvenet_amd import HSC, BatchSystem, CutLaplacian, meshgen
    laps = [CutLaplacian.from_mesh(meshgen.capsule_union(c, c + 2, t, seed=11 + c))
            for c, t in ((1, 3000), (2, 9000), (3, 20000))]
    batch = BatchSystem(laps)
    rng = np.random.default_rng(7)
    B = batch.stack([l.harmonic_rhs(rng.uniform(-0.2, 0.2, (l.n_con, 3))) for l in laps])
    h = HSC(batch, max_levels=48)
    x0, i0 = batch.solve(B, precond="hsc", precision="mixed", tol=1e-9, hsc=h, graph=False)
    x1, i1 = batch.solve(B, precond="hsc", precision="mixed", tol=1e-9, hsc=h, graph=True)
    assert i0.status == 0 and i1.iters == i0.iters
    assert np.array_equal(x0.cpu().numpy(), x1.cpu().numpy())


def test_autotune_choice_converges_and_caches(gpu, tmp_path):
    """The autotuner tries the launch shapes, keeps one whose solve reaches
    the tolerance, stores it, and a second call takes it from the cache."""
    import torch
    from tool_godot_curvenet_amd import HSC, CutLaplacian, autotune, meshgen
    lap = CutLaplacian.from_mesh(meshgen.config_mesh("icosphere6"))
    B = torch.as_tensor(lap.harmonic_rhs(meshgen.seeded_boundary(lap.n_con, 64, seed=9)),
                        device="cuda")
    h = HSC(lap)
    path = str(tmp_path / "tune.json")
    r = autotune.autotune(lap, B, precond="hsc", precision="mixed", hsc=h, path=path, reps=1)
    assert not r["cached"] and len(r["candidates"]) >= 4
    assert lap.get_tune() == r["tune"]
    assert r["inner_tol"] in autotune.INNER_TOLS and lap.inner_tol == r["inner_tol"]
    assert {c["inner_tol"] for c in r["candidates"]} == set(autotune.INNER_TOLS)
    x, info = lap.solve(B, precond="hsc", precision="mixed", tol=1e-9, hsc=h)
    assert info.status == 0
    A = lap.A_ff()
    Bh = B.cpu().numpy()
    R = Bh - O.spmm(A.row_ptr, A.col_idx, A.values, x.cpu().numpy())
    assert np.all(np.linalg.norm(R, axis=0) <= 1e-9 * np.linalg.norm(Bh, axis=0))
    lap.set_tune(None)
    lap.inner_tol = lap.INNER_TOL
    r2 = autotune.autotune(lap, B, precond="hsc", precision="mixed", hsc=h, path=path)
    assert r2["cached"] and r2["tune"] == r["tune"] and lap.get_tune() == r["tune"]
    assert lap.inner_tol == r["inner_tol"]
    # another right-hand side of the same system: only the drop is picked,
    # the launch shapes stay
    r3 = autotune.autotune(lap, 2.0 * B + 1.0, precond="hsc", precision="mixed", hsc=h,
                           path=path, shapes=False, reps=1)
    assert not r3["cached"] and r3["tune"] == r["tune"] and lap.get_tune() == r["tune"]
    assert r3["inner_tol"] in autotune.INNER_TOLS and lap.inner_tol == r3["inner_tol"]
    assert {c["inner_tol"] for c in r3["candidates"]} == set(autotune.INNER_TOLS)


@pytest.mark.parametrize("inner_tol", [0.0316, 1e-2, 1e-3, 1e-4])
def test_inner_tol_converges(gpu, inner_tol):
    """Every residual-replacement drop the autotuner may pick reaches 1e-9 in
    the FP64 residual recomputed on the CPU; a larger drop takes fewer
    replacements."""
    import torch
    from tool_godot_curvenet_amd import HSC, CutLaplacian, meshgen
    lap = CutLaplacian.from_mesh(meshgen.config_mesh("icosphere6"))
    B = torch.as_tensor(lap.harmonic_rhs(meshgen.seeded_boundary(lap.n_con, 48, seed=4)),
                        device="cuda")
    h = HSC(lap)
    x, info = lap.solve(B, precond="hsc", precision="mixed", tol=1e-9, hsc=h,
                        inner_tol=inner_tol)
    x0, info0 = lap.solve(B, precond="hsc", precision="mixed", tol=1e-9, hsc=h,
                          inner_tol=0.0316)
    assert info.status == 0
    A = lap.A_ff()
    Bh = B.cpu().numpy()
    R = Bh - O.spmm(A.row_ptr, A.col_idx, A.values, x.cpu().numpy())
    assert np.all(np.linalg.norm(R, axis=0) <= 1e-9 * np.linalg.norm(Bh, axis=0))
    assert info.refinements <= info0.refinements


def test_inner_tol_out_of_range_refused(gpu):
    import ctypes
    import torch
    from tool_godot_curvenet_amd import CutLaplacian, _lib, meshgen
    lap = CutLaplacian.from_mesh(meshgen.config_mesh("cylinder64"))
    lib = _lib.load()
    B = torch.ones((lap.n_free, 3), dtype=torch.float64, device="cuda")
    X = torch.zeros_like(B)
    for bad in (0.0, 1.0, -0.5):
        o = _lib.PcgOpts()
        lib.cn_pcg_default_opts(ctypes.byref(o))
        o.precond = _lib.CN_PRECOND_JACOBI
        o.inner_tol = bad
        info = _lib.SolveInfo()
        st = lib.cn_pcg_solve(lap.handle, ctypes.c_void_p(0), ctypes.byref(o), 3,
                              ctypes.c_void_p(B.data_ptr()), ctypes.c_void_p(X.data_ptr()),
                              ctypes.byref(info), ctypes.c_void_p(0))
        assert st == _lib.CN_BAD_ARG


@pytest.mark.parametrize("precond,precision,tol", [("hsc", "fp64", 1e-12), ("hsc", "mixed", 1e-5),
                                                   ("sgs", "fp64", 1e-12), ("sgs", "mixed", 1e-5)])
def test_fused_iteration_equals_unfused(gpu, precond, precision, tol):
    """The fused iteration (direction update on the product, residual and
    iterate updates on a second product) computes the unfused path's
    vectors with the same arithmetic; only the partial sums of p.q and r.r
    are grouped differently.  Three iterations agree to rounding, and both
    solves converge in the same number of iterations (within one)."""
    from tool_godot_curvenet_amd import HSC, SGS, CutLaplacian, meshgen
    lap = CutLaplacian.from_mesh(meshgen.config_mesh("icosphere6"))
    k = 64
    B = lap.harmonic_rhs(meshgen.seeded_boundary(lap.n_con, k, seed=21))
    kw = {"hsc": HSC(lap)} if precond == "hsc" else {"sgs": SGS(lap)}
    out = {}
    for form in (1, 2):
        lap.set_tune({"spmm_form": form})
        x3, i3 = lap.solve(B, precond=precond, precision=precision, tol=1e-9, max_iter=3, **kw)
        xf, inf = lap.solve(B, precond=precond, precision=precision, tol=1e-9, **kw)
        out[form] = (x3.cpu().numpy(), xf.cpu().numpy(), i3, inf)
    lap.set_tune(None)
    a, b = out[1], out[2]
    assert np.abs(a[0] - b[0]).max() <= tol * np.abs(a[0]).max()
    assert a[3].status == 0 and b[3].status == 0 and abs(a[3].iters - b[3].iters) <= 1
    A = lap.A_ff()
    R = B - O.spmm(A.row_ptr, A.col_idx, A.values, b[1])
    assert np.all(np.linalg.norm(R, axis=0) <= 1e-9 * np.linalg.norm(B, axis=0))


def test_fused_graph_odd_block(gpu):
    """SGS polls every check_every iterations: with 3 the captured block
    flips the two P buffers; replays from either parity equal direct
    launches bit for bit."""
    from tool_godot_curvenet_amd import SGS, CutLaplacian, meshgen
    lap = CutLaplacian.from_mesh(meshgen.config_mesh("icosphere6"))
    B = lap.harmonic_rhs(meshgen.seeded_boundary(lap.n_con, 32, seed=23))
    g = SGS(lap)
    x0, i0 = lap.solve(B, precond="sgs", precision="mixed", tol=1e-9, sgs=g, check_every=3,
                       graph=False)
    for _ in range(2):
        x1, i1 = lap.solve(B, precond="sgs", precision="mixed", tol=1e-9, sgs=g, check_every=3,
                           graph=True)
        assert i1.iters == i0.iters and np.array_equal(x0.cpu().numpy(), x1.cpu().numpy())


# Column scales: from x0 = 0 a column whose ||b|| lies within [1e-15, 1e15]
# runs unscaled, others at 1/||b||; a solve with columns of both kinds
# scales every column by its own rule.  Every column converges to the
# tolerance relative to its own ||b||, and its solution divided by its scale
# is the unit-scale column's solution to what that tolerance allows (the
# columns' FP32 roundings differ, so their iterates do).
@pytest.mark.parametrize("precond", ["hsc", "jacobi"])
@pytest.mark.parametrize("scales", [
    (1.0, 1e-12, 1e12, 1e-14, 1e14, 3.0, 1e-7, 1e7),       # all at scale 1
    (1.0, 1e-25, 1e25, 1e-18, 1e18, 1e-40, 1e40, 1e5),     # mixed rules
])
def test_pcg_column_scales(gpu, precond, scales):
    from tool_godot_curvenet_amd import HSC
    from tool_godot_curvenet_amd import meshgen as G
    lap = _lap("icosphere6")
    A = lap.A_ff()
    k = len(scales)
    B0 = lap.harmonic_rhs(G.seeded_boundary(lap.n_con, 1, seed=77))
    B = B0 * np.asarray(scales)[None, :]
    kw = {"hsc": HSC(lap)} if precond == "hsc" else {}
    x, info = lap.solve(B, precond=precond, precision="mixed", tol=1e-9, **kw)
    assert info.status == 0
    X = x.cpu().numpy()
    R = B - O.spmm(A.row_ptr, A.col_idx, A.values, X)
    rr = np.linalg.norm(R, axis=0) / np.linalg.norm(B, axis=0)
    assert np.all(rr <= 1e-9), rr
    assert np.all(np.isfinite(X))
    # every column is the same system: X[:, j] / scale_j agrees with column 0
    X1 = X / np.asarray(scales)[None, :]
    dev = np.linalg.norm(X1 - X1[:, :1], axis=0) / np.linalg.norm(X1[:, 0])
    assert np.all(dev <= 1e-4), dev
