"""The safetensors container (host code in the library, no GPU) against the
reference's own test vectors (tests/safetensors/vectors, copied under
tests/golden/safetensors), an independent byte-level parse, the `safetensors`
package, and the LeanSafetensors rules for malformed files; Mire mesh ingest
and the weld."""
import json
import pathlib
import struct

import numpy as np
import pytest

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden" / "safetensors"
VECTORS = sorted(GOLDEN.glob("*.safetensors"))


def _F():
    from tool_godot_curvenet_amd import formats
    return formats


def _raw(data: bytes):
    """Byte-level parse keeping the file's key order: (tensors, metadata)."""
    hlen = struct.unpack_from("<Q", data, 0)[0]
    header = json.loads(data[8:8 + hlen])
    body = data[8 + hlen:]
    meta = header.pop("__metadata__", {}) or {}
    return {k: (v["dtype"], tuple(v["shape"]), body[v["data_offsets"][0]:v["data_offsets"][1]])
            for k, v in header.items()}, meta


def test_vectors_present():
    assert len(VECTORS) == 9


@pytest.mark.parametrize("path", VECTORS, ids=lambda p: p.stem)
def test_reference_vectors_parse(path):
    st = _F().read_safetensors(path)
    raw, meta = _raw(path.read_bytes())
    assert list(st.tensors) == sorted(raw)        # key-sorted, as the reference decodes
    for k, (tag, shape, data) in raw.items():
        t, s, a = st.tensors[k]
        assert (t, s) == (tag, shape)
        assert a.tobytes() == data
    assert st.metadata == meta


@pytest.mark.parametrize("path", VECTORS, ids=lambda p: p.stem)
def test_reference_vectors_against_safetensors_package(path):
    from safetensors.numpy import load_file
    st = _F().read_safetensors(path)
    if any(t == "BF16" for t, _, _ in st.tensors.values()):
        # numpy has no bfloat16: the package cannot materialise the file
        return
    ref = load_file(str(path))
    assert set(ref) == set(st.tensors)
    for k, a in ref.items():
        assert np.array_equal(st.get(k), a) and st.get(k).dtype == a.dtype


@pytest.mark.parametrize("path", VECTORS, ids=lambda p: p.stem)
def test_writer_reproduces_vector_bytes(path):
    """Written in the file's own order (metadata first, tensors by offset),
    the encoder gives the vector's bytes: same compact JSON, same 8-byte
    space padding."""
    data = path.read_bytes()
    raw, meta = _raw(data)
    hdr = json.loads(data[8:8 + struct.unpack_from("<Q", data, 0)[0]])
    order = sorted((k for k in hdr if k != "__metadata__"), key=lambda k: hdr[k]["data_offsets"])
    st = _F().read_safetensors(path)
    tensors = [(k, (st.tensors[k][0], st.tensors[k][2])) for k in order]
    assert _F().encode_safetensors(tensors, meta) == data


def test_round_trip_every_dtype(tmp_path):
    from safetensors.numpy import load_file
    F = _F()
    rng = np.random.default_rng(0)
    t = {"a/f64": rng.normal(size=(3, 4)), "b.f32": rng.normal(size=5).astype(np.float32),
         "c f16": np.arange(6, dtype=np.float16).reshape(2, 3), "d": np.array(7, np.int64),
         "e\"quote\\": np.arange(4, dtype=np.uint16), "empty": np.zeros((0, 3), np.int32),
         "flags": np.array([True, False, True]), "snow☃": np.arange(3, dtype=np.int8)}
    p = tmp_path / "x.safetensors"
    F.write_safetensors(p, t, {"k\n": "v\té", "fmt": "x"})
    st = F.read_safetensors(p)
    for k, a in t.items():
        assert st.get(k).dtype == a.dtype and np.array_equal(st.get(k), a), k
    assert st.metadata == {"k\n": "v\té", "fmt": "x"}
    ref = load_file(str(p))
    for k, a in t.items():
        assert np.array_equal(ref[k], a), k
    # BF16 bits pass through unchanged
    bits = np.array([0x3F80, 0xC000, 0x7F80], np.uint16)
    st = F.parse_safetensors(F.encode_safetensors([("h", ("BF16", bits))]))
    assert st.tensors["h"][0] == "BF16" and np.array_equal(st.get("h"), bits)
    # the body starts on an 8-byte boundary
    data = p.read_bytes()
    assert (8 + struct.unpack_from("<Q", data, 0)[0]) % 8 == 0


def _file(header: dict, body: bytes = b"", raw_header: bytes | None = None) -> bytes:
    h = raw_header if raw_header is not None else json.dumps(header).encode()
    return struct.pack("<Q", len(h)) + h + body


@pytest.mark.parametrize("data,msg", [
    (b"\x01\x02", "< 8 bytes"),
    (struct.pack("<Q", 100) + b"{}", "exceeds file size"),
    (_file({"x": {"dtype": "F32", "shape": [2], "data_offsets": [0, 8]}}, b"\0" * 4),
     "beyond body"),
    (_file({"x": {"dtype": "F32", "shape": [3], "data_offsets": [0, 8]}}, b"\0" * 8),
     "dtype"),
    (_file({"x": {"dtype": "F32", "shape": [1], "data_offsets": [8, 4]}}, b"\0" * 8), "lo > hi"),
    (_file({"x": {"dtype": "Q8", "shape": [1], "data_offsets": [0, 1]}}, b"\0"),
     "unknown dtype"),
    (_file({"x": {"dtype": "U8", "shape": [1.0], "data_offsets": [0, 1]}}, b"\0"),
     "shape entry"),
    (_file({"x": {"dtype": "U8", "shape": [-1], "data_offsets": [0, 1]}}, b"\0"),
     "shape entry"),
    (_file({"x": {"dtype": "U8", "shape": [1], "data_offsets": [0]}}, b"\0"), "2 entries"),
    (_file({"x": {"dtype": "U8", "shape": [1]}}, b"\0"), "data_offsets"),
    (_file({"x": [1, 2]}), "not an object"),
    (_file({}, raw_header=b"[1]"), "root is not an object"),
    (_file({}, raw_header=b'{"a": tru}'), "JSON"),
    (_file({"__metadata__": [1]}), "__metadata__"),
])
def test_malformed_files_rejected(data, msg):
    from tool_godot_curvenet_amd._lib import CurvenetError
    with pytest.raises(CurvenetError, match=msg):
        _F().parse_safetensors(data)


def test_header_decoding_rules():
    """Key order, last duplicate, non-string metadata as JSON text, escapes."""
    hdr = (b'{"__metadata__":{"n":3,"s":"a\\u00e9\\ud83d\\ude00","l":[1,{"x":null}]},'
           b'"b":{"dtype":"U8","shape":[1],"data_offsets":[0,1]},'
           b'"a":{"dtype":"U8","shape":[1],"data_offsets":[0,1]},'
           b'"a":{"dtype":"U8","shape":[1],"data_offsets":[1,2]}}')
    st = _F().parse_safetensors(_file({}, b"\x05\x06", raw_header=hdr))
    assert list(st.tensors) == ["a", "b"]
    assert st.get("a")[0] == 6 and st.get("b")[0] == 5       # the later "a" wins
    assert st.metadata == {"l": '[1,{"x":null}]', "n": "3", "s": "aé\U0001F600"}


# ------------------------------------------------------------- mesh ingest --
def _expanded(pos, tris):
    """The exporter's per-loop expansion: one position per triangle corner."""
    loops = np.asarray(pos, np.float32)[np.asarray(tris).reshape(-1)]
    return loops, np.arange(len(loops), dtype=np.uint32).reshape(-1, 3)


def test_weld_recovers_the_mesh():
    from tool_godot_curvenet_amd import meshgen
    m = meshgen.config_mesh("cylinder64")
    P, I = _expanded(m.pos, m.tris)
    pos, tris, lv = _F().weld(P, I)
    assert pos.shape[0] == m.pos.shape[0] and tris.shape == m.tris.shape
    # first-appearance numbering: relabel the original the same way
    first = {}
    for v in np.asarray(m.tris).reshape(-1):
        first.setdefault(int(v), len(first))
    relabel = np.array([first[v] for v in range(len(first))])
    assert np.array_equal(tris, relabel[np.asarray(m.tris)])
    assert np.array_equal(pos[relabel], np.asarray(m.pos, np.float32).astype(np.float64))
    assert np.array_equal(lv, relabel[np.asarray(m.tris).reshape(-1)])


def test_load_mire_file(tmp_path):
    from tool_godot_curvenet_amd import meshgen
    F = _F()
    m = meshgen.config_mesh("cylinder64")
    P, I = _expanded(m.pos, m.tris)
    n = len(P)
    names = ["Body.001", "arm [L]"]
    t, meta = {}, {"format": "curvenet-mesh-v1", "source_axes": "Z-up",
                   "mesh_names": json.dumps(names), "mesh_n_verts": json.dumps([n, 3]),
                   "mesh_n_tris": json.dumps([len(I), 1])}
    t[f"mesh.{F.safe_name(names[0])}.positions"] = P
    t[f"mesh.{F.safe_name(names[0])}.indices"] = I
    t[f"mesh.{F.safe_name(names[0])}.weights"] = np.full((n, 4), 0.25, np.float32)
    t[f"mesh.{F.safe_name(names[0])}.joints"] = np.zeros((n, 4), np.uint16)
    t[f"mesh.{F.safe_name(names[1])}.positions"] = np.eye(3, dtype=np.float32)
    t[f"mesh.{F.safe_name(names[1])}.indices"] = np.array([[0, 1, 2]], np.uint32)
    p = tmp_path / "mire.safetensors"
    F.write_safetensors(p, t, meta)
    meshes = F.load_mire(p)
    assert list(meshes) == names
    body = meshes["Body.001"]
    assert body["pos"].shape == (m.pos.shape[0], 3) and body["tris"].shape == m.tris.shape
    assert body["weights"].shape == (n, 4) and body["joints"].dtype == np.uint16
    assert meshes["arm [L]"]["tris"].tolist() == [[0, 1, 2]]
    # an index past the positions is an error, not a silent read
    t[f"mesh.{F.safe_name(names[1])}.indices"] = np.array([[0, 1, 3]], np.uint32)
    F.write_safetensors(p, t, meta)
    with pytest.raises(ValueError, match="index past"):
        F.load_mire(p)


def test_reference_mesh_like_vector():
    """The reference's mesh_like vector: 4 vertices, 2 triangles, weights."""
    st = _F().read_safetensors(GOLDEN / "mesh_like.safetensors")
    P, I = st.get("positions"), st.get("indices")
    pos, tris, _ = _F().weld(P, I)
    assert pos.shape == (4, 3) and tris.shape == (2, 3)
    assert st.metadata["n_verts"] == "4" and st.metadata["n_tris"] == "2"
    assert np.allclose(st.get("weights").sum(axis=1), 1.0)


class _FakeSystem:
    """Enough of a solver system for the sidecar checks (no device): a
    4-row CSR and the vertex count."""

    def __init__(self):
        from tool_godot_curvenet_amd.curvenet import CsrMatrix
        self._A = CsrMatrix(4, 4, np.array([0, 1, 2, 3, 4], np.int32),
                            np.arange(4, dtype=np.int32), np.full(4, 2.0))
        self.n_free = 4

        class _Info:
            nv = 4
        self.info = _Info()

    def A_ff(self):
        return self._A


def _sidecar(tmp_path, system, tensors, levels=0):
    F = _F()
    meta = {"format": F.SIDECAR_FORMAT, "fingerprint": F.system_fingerprint(system),
            "n_free": "4", "hsc_levels": str(levels), "hsc_theta": "0.015",
            "hsc_coarse_max": "512", "hsc_max_levels": "24", "hsc_min_frac": "0.05",
            "hsc_compensate": "0"}
    path = tmp_path / "side.safetensors"
    F.write_safetensors(str(path), tensors, meta)
    return str(path)


@pytest.mark.parametrize("bad", ["coarse_L", "coarse_inv"])
@pytest.mark.parametrize("dtype", [np.float32, np.int32, np.uint8])
def test_load_hsc_rejects_non_f64_coarse(tmp_path, bad, dtype):
    """cn_hsc_import reads n_coarse^2 doubles from each coarse matrix: any
    other dtype is refused before the library is called."""
    sysm = _FakeSystem()
    t = {"hsc.coarse_L": np.eye(4), "hsc.coarse_inv": np.eye(4)}
    t["hsc." + bad] = np.eye(4).astype(dtype)
    with pytest.raises(ValueError, match="F64"):
        _F().load_hsc(_sidecar(tmp_path, sysm, t), sysm)


@pytest.mark.parametrize("start,nidx,msg", [
    ([0, 1, 2, 3], 3, "entries"),                 # nv + 1 = 5 expected
    ([1, 1, 2, 3, 4], 4, "begin at 0"),
    ([0, 2, 1, 3, 4], 4, "never decrease"),
    ([0, 1, 2, 3, 5], 4, "differ"),
])
def test_influences_inconsistent_csr_rejected(tmp_path, start, nidx, msg):
    F = _F()
    with pytest.raises(ValueError, match=msg):
        F.check_influences(4, np.array(start, np.int32), np.zeros(nidx, np.int32),
                           np.zeros(nidx))
    sysm = _FakeSystem()
    t = {"ddm.start": np.array(start, np.int32), "ddm.idx": np.zeros(nidx, np.int32),
         "ddm.w": np.zeros(nidx)}
    with pytest.raises(ValueError, match=msg):
        F.load_influences(_sidecar(tmp_path, sysm, t), sysm)


def test_influences_dtypes_checked(tmp_path):
    sysm = _FakeSystem()
    t = {"ddm.start": np.arange(5, dtype=np.int64), "ddm.idx": np.zeros(4, np.int32),
         "ddm.w": np.zeros(4)}
    with pytest.raises(ValueError, match="I32"):
        _F().load_influences(_sidecar(tmp_path, sysm, t), sysm)
    t["ddm.start"] = np.arange(5, dtype=np.int32)
    s, i, w = _F().load_influences(_sidecar(tmp_path, sysm, t), sysm)
    assert list(s) == [0, 1, 2, 3, 4]
