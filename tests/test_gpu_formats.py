"""The bind-cache sidecar on the GPU: a hierarchy written to safetensors and
rebuilt with cn_hsc_import applies and solves bit-identically; sidecars of
another system or with broken structure are refused; skinning influences
round-trip; a mesh ingested from a Mire-style file solves."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _lap(name):
    from tool_godot_curvenet_amd import CutLaplacian, meshgen
    return CutLaplacian.from_mesh(meshgen.config_mesh(name))


@pytest.mark.parametrize("name", ["cylinder64", "icosphere6"])
def test_hsc_sidecar_round_trip(gpu, tmp_path, name):
    from tool_godot_curvenet_amd import HSC
    from tool_godot_curvenet_amd import meshgen as G
    lap = _lap(name)
    H = HSC(lap)
    p = tmp_path / "cache.safetensors"
    H.save(p)
    H2 = HSC.load(lap, p)
    assert H2.info.nlevels == H.info.nlevels and H2.info.n_coarse == H.info.n_coarse
    assert H2.info.nnz == H.info.nnz
    r = np.random.default_rng(1).normal(size=(lap.n_free, 7))
    assert np.array_equal(H2.apply(r), H.apply(r))
    for a, b in zip(H.levels(), H2.levels()):
        for k in ("F", "C", "dinv"):
            assert np.array_equal(a[k], b[k])
        for k in ("R", "P"):
            for x, y in zip(a[k], b[k]):
                assert np.array_equal(x, y)
    B = lap.harmonic_rhs(G.seeded_boundary(lap.n_con, 3, seed=9))
    x1, i1 = lap.solve(B, precond="hsc", precision="mixed", tol=1e-9, hsc=H)
    x2, i2 = lap.solve(B, precond="hsc", precision="mixed", tol=1e-9, hsc=H2)
    assert i1.iters == i2.iters and i1.status == i2.status == 0
    assert np.array_equal(x1.cpu().numpy(), x2.cpu().numpy())


def test_sidecar_of_another_system_refused(gpu, tmp_path):
    from tool_godot_curvenet_amd import HSC
    lap = _lap("cylinder64")
    p = tmp_path / "cache.safetensors"
    HSC(lap).save(p)
    other = _lap("cylinder64")
    pos = other.pos.copy()
    pos[5] += 1e-3
    other.set_positions(pos)                 # same pattern, other values
    with pytest.raises(ValueError, match="another system"):
        HSC.load(other, p)


def test_broken_hierarchy_refused(gpu, tmp_path):
    from tool_godot_curvenet_amd import HSC, formats
    from tool_godot_curvenet_amd._lib import CurvenetError
    lap = _lap("cylinder64")
    p = tmp_path / "cache.safetensors"
    HSC(lap).save(p)
    st = formats.read_safetensors(p)
    for key, fix in (("hsc.00.F", lambda a: a[::-1].copy()),            # not ascending
                     ("hsc.00.p_col", lambda a: a + 10 ** 6),           # past the next level
                     ("hsc.00.r_col", lambda a: np.full_like(a, -1))):  # not an F row
        t = {k: v[2] for k, v in st.tensors.items()}
        t[key] = fix(t[key])
        q = tmp_path / "bad.safetensors"
        formats.write_safetensors(q, t, st.metadata)
        with pytest.raises(CurvenetError, match="hsc_import"):
            HSC.load(lap, q)


def test_influences_sidecar(gpu, tmp_path):
    from tool_godot_curvenet_amd import ddm, formats
    lap = _lap("cylinder64")
    start, idx, w, _ = ddm.bind(lap, lap.tris, nu=1, K=4, precision="fp64", tol=1e-10)
    p = tmp_path / "cache.safetensors"
    formats.save_sidecar(p, lap, influences=(start, idx, w))
    s2, i2, w2 = formats.load_influences(p, lap)
    assert np.array_equal(s2, start.cpu().numpy()) and np.array_equal(i2, idx.cpu().numpy())
    assert np.array_equal(w2, w.cpu().numpy())


def test_mire_ingest_solves(gpu, tmp_path):
    """A mesh written in the exporter's per-loop layout, ingested (welded),
    cut along a ring and solved with HSC; the residual checked by the oracle."""
    from tool_godot_curvenet_amd import CutLaplacian, cut_seams, formats, meshgen
    from tool_godot_curvenet_amd import meshgen as G
    m = meshgen.config_mesh("cylinder64")
    loops = np.asarray(m.pos, np.float32)[np.asarray(m.tris).reshape(-1)]
    idx = np.arange(len(loops), dtype=np.uint32).reshape(-1, 3)
    p = tmp_path / "mire.safetensors"
    formats.write_safetensors(p, {"mesh.cyl.positions": loops, "mesh.cyl.indices": idx},
                              {"format": "curvenet-mesh-v1", "mesh_names": '["cyl"]'})
    mesh = formats.load_mire(p)["cyl"]
    # the seam edges, renumbered by the weld
    first = {}
    for v in np.asarray(m.tris).reshape(-1):
        first.setdefault(int(v), len(first))
    cut = np.array([[first[int(a)], first[int(b)]] for a, b in m.cut_edges], np.int32)
    nt, origin, con, _ = cut_seams(len(mesh["pos"]), mesh["tris"], cut)
    lap = CutLaplacian(mesh["pos"][origin], nt, con)
    assert lap.n_con == _lap("cylinder64").n_con
    A = lap.A_ff()
    B = lap.harmonic_rhs(G.seeded_boundary(lap.n_con, 3, seed=4))
    x, info = lap.solve(B, precond="hsc", precision="mixed", tol=1e-9)
    assert info.status == 0
    R = B - O.spmm(A.row_ptr, A.col_idx, A.values, x.cpu().numpy())
    assert np.all(np.linalg.norm(R, axis=0) <= 1e-9 * np.linalg.norm(B, axis=0))
