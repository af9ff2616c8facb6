"""On-disk formats: the safetensors container (LeanSafetensors), ingest of
Mire-style mesh files (SafetensorsToGltf.lean:11-36, written by
misc/blend_to_safetensors.py), and the hierarchy / skinning sidecar that
lets a session skip the bind (Features/KnownUnknowns/BindCachePersistence.lean:3-5).

The container code is the library's (``cn_st_*`` in ``csrc/cn_safetensors.cpp``);
this module turns its tensors into numpy arrays.  None of it needs a GPU
except building and loading a hierarchy.
"""
from __future__ import annotations

import ctypes as C
import hashlib
import json

import numpy as np

from . import _lib as L

# dtype codes of the C ABI (Header.lean:36-41) and their numpy forms; BF16
# has no numpy type and comes back as its raw uint16 bits
DTYPES = ["F32", "F64", "F16", "BF16", "U8", "U16", "U32", "U64", "I8", "I16", "I32", "I64",
          "BOOL"]
_NP = {"F32": np.float32, "F64": np.float64, "F16": np.float16, "BF16": np.uint16,
       "U8": np.uint8, "U16": np.uint16, "U32": np.uint32, "U64": np.uint64, "I8": np.int8,
       "I16": np.int16, "I32": np.int32, "I64": np.int64, "BOOL": np.bool_}
_CODE = {np.dtype(np.float32): 0, np.dtype(np.float64): 1, np.dtype(np.float16): 2,
         np.dtype(np.uint8): 4, np.dtype(np.uint16): 5, np.dtype(np.uint32): 6,
         np.dtype(np.uint64): 7, np.dtype(np.int8): 8, np.dtype(np.int16): 9,
         np.dtype(np.int32): 10, np.dtype(np.int64): 11, np.dtype(np.bool_): 12}


class SafeTensors:
    """A parsed safetensors file (SafeTensors.parse + validate, Reader.lean:33-61).
    ``tensors`` maps name -> (dtype tag, shape, numpy array) in key order;
    ``metadata`` maps key -> string."""

    def __init__(self, handle):
        lib = L.load()
        self.tensors: dict = {}
        self.metadata: dict = {}
        try:
            for i in range(lib.cn_st_count(handle)):
                t = L.StTensor()
                L.check(lib.cn_st_tensor_at(handle, i, C.byref(t)), "st_tensor_at")
                tag = DTYPES[t.dtype]
                shape = tuple(t.shape[d] for d in range(t.ndim))
                raw = C.string_at(t.data, t.nbytes) if t.nbytes else b""
                arr = np.frombuffer(raw, dtype=np.dtype(_NP[tag]).newbyteorder("<"))
                self.tensors[t.name.decode()] = (tag, shape, arr.reshape(shape).copy())
            for i in range(lib.cn_st_metadata_count(handle)):
                k, v = C.c_char_p(), C.c_char_p()
                L.check(lib.cn_st_metadata_at(handle, i, C.byref(k), C.byref(v)),
                        "st_metadata_at")
                self.metadata[k.value.decode()] = v.value.decode()
        finally:
            lib.cn_st_free(handle)

    def get(self, name: str):
        """SafeTensors.get (Reader.lean:64-69): the array, or None."""
        t = self.tensors.get(name)
        return None if t is None else t[2]


def read_safetensors(path) -> SafeTensors:
    h = C.c_void_p(0)
    L.check(L.load().cn_st_read_file(str(path).encode(), C.byref(h)), "st_read_file")
    return SafeTensors(h)


def parse_safetensors(data: bytes) -> SafeTensors:
    h = C.c_void_p(0)
    buf = C.create_string_buffer(data, len(data))
    L.check(L.load().cn_st_parse(buf, len(data), C.byref(h)), "st_parse")
    return SafeTensors(h)


def _entries(tensors):
    """(StTensor array, keep-alive list) for [(name, array or (tag, array))]."""
    items = list(tensors.items()) if isinstance(tensors, dict) else list(tensors)
    arr = (L.StTensor * max(1, len(items)))()
    keep = []
    for i, (name, val) in enumerate(items):
        if isinstance(val, tuple):
            tag, a = val
            code = DTYPES.index(tag)
        else:
            a = val
            code = _CODE[np.dtype(a.dtype)]
        a = np.asarray(a)
        a = a.astype(a.dtype.newbyteorder("<"), order="C", copy=False)  # keeps 0-d shapes
        shape = (C.c_int64 * max(1, a.ndim))(*a.shape)
        nm = name.encode()
        keep += [a, shape, nm]
        arr[i].name = nm
        arr[i].dtype = code
        arr[i].ndim = a.ndim
        arr[i].shape = C.cast(shape, C.POINTER(C.c_int64))
        arr[i].data = a.ctypes.data if a.nbytes else None
        arr[i].nbytes = a.nbytes
    return arr, len(items), keep


def _meta(metadata):
    md = list((metadata or {}).items())
    keys = (C.c_char_p * max(1, len(md)))(*[k.encode() for k, _ in md])
    vals = (C.c_char_p * max(1, len(md)))(*[str(v).encode() for _, v in md])
    return keys, vals, len(md)


def encode_safetensors(tensors, metadata=None) -> bytes:
    """Writer.build (Writer.lean:54-58): tensors in the given order (a dict or
    (name, array) pairs; an array may be given as (dtype tag, array) to write
    BF16 bits), metadata strings first."""
    lib = L.load()
    arr, n, _keep = _entries(tensors)
    keys, vals, nm = _meta(metadata)
    size = C.c_int64(0)
    L.check(lib.cn_st_encode(n, arr, nm, keys, vals, None, 0, C.byref(size)), "st_encode")
    buf = C.create_string_buffer(size.value)
    L.check(lib.cn_st_encode(n, arr, nm, keys, vals, buf, size.value, C.byref(size)),
            "st_encode")
    return buf.raw


def write_safetensors(path, tensors, metadata=None) -> None:
    arr, n, _keep = _entries(tensors)
    keys, vals, nm = _meta(metadata)
    L.check(L.load().cn_st_write_file(str(path).encode(), n, arr, nm, keys, vals),
            "st_write_file")


# ------------------------------------------------------------- mesh ingest --
def safe_name(s: str) -> str:
    """The exporter's tensor-name normalisation (misc/blend_to_safetensors.py
    _safe_name): '.', ' ', '/', '[' and ']' become '_'."""
    for c in ". /[]":
        s = s.replace(c, "_")
    return s


def weld(positions, indices):
    """Merge the per-loop expanded vertices of a Mire mesh into shared
    vertices: equal FP32 position triples (bit patterns) become one vertex,
    numbered in order of first appearance.  Returns (pos nv x 3 FP64,
    tris nf x 3 int32, loop_vertex: the welded vertex of every loop vertex)."""
    p = np.ascontiguousarray(np.asarray(positions, np.float32).reshape(-1, 3))
    rows = p.view(np.dtype((np.void, 12))).reshape(-1)
    _, first, inv = np.unique(rows, return_index=True, return_inverse=True)
    order = np.argsort(first, kind="stable")            # unique rows by first appearance
    rank = np.empty_like(order)
    rank[order] = np.arange(len(order))
    loop_vertex = rank[inv.reshape(-1)].astype(np.int32)
    pos = p[first[order]].astype(np.float64)
    tris = loop_vertex[np.asarray(indices, np.int64).reshape(-1, 3)].astype(np.int32)
    return pos, tris, loop_vertex


def load_mire(path) -> dict:
    """Meshes of a Mire-style file (`format` "curvenet-mesh-v1"): name ->
    dict(pos, tris (welded), loop_vertex, joints, weights, normals, uvs as
    stored per loop vertex)."""
    st = read_safetensors(path)
    names = json.loads(st.metadata.get("mesh_names", "[]"))
    if not names:  # files without the metadata: every mesh.<name>.positions
        names = sorted({k.split(".")[1] for k in st.tensors
                        if k.startswith("mesh.") and k.endswith(".positions")})
    out = {}
    for nm in names:
        pre = f"mesh.{safe_name(nm)}."
        P, I = st.get(pre + "positions"), st.get(pre + "indices")
        if P is None or I is None:
            raise ValueError(f"{path}: mesh {nm!r} lacks positions or indices")
        if P.ndim != 2 or P.shape[1] != 3 or I.ndim != 2 or I.shape[1] != 3:
            raise ValueError(f"{path}: mesh {nm!r}: positions must be N x 3, indices M x 3")
        if I.size and (int(I.max()) >= len(P)):
            raise ValueError(f"{path}: mesh {nm!r}: index past the positions")
        pos, tris, lv = weld(P, I)
        out[nm] = {"pos": pos, "tris": tris, "loop_vertex": lv,
                   **{k: st.get(pre + k) for k in ("joints", "weights", "normals", "uvs")}}
    return out


# ------------------------------------------------------------- sidecar ------
SIDECAR_FORMAT = "curvenet-cache-v1"


def system_fingerprint(system) -> str:
    """SHA-256 of the system's A_ff (row pointers, columns, FP64 values): a
    hierarchy belongs to the values it was built from."""
    A = system.A_ff()
    h = hashlib.sha256()
    for a in (np.int64(A.rows), A.row_ptr, A.col_idx, A.values):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def save_sidecar(path, system, hsc=None, influences=None) -> None:
    """Write the bind results of `system` to a safetensors sidecar: the HSC
    hierarchy (every level's F, C, 1/a_ff, R, P and the coarse factor and
    inverse) and/or skinning influences (start, idx, w), with the system's
    fingerprint."""
    t = {}
    meta = {"format": SIDECAR_FORMAT, "fingerprint": system_fingerprint(system),
            "n_free": str(system.n_free)}
    if hsc is not None:
        o = hsc.opts
        meta.update({"hsc_levels": str(hsc.info.nlevels), "hsc_theta": repr(o.theta),
                     "hsc_coarse_max": str(o.coarse_max), "hsc_max_levels": str(o.max_levels),
                     "hsc_min_frac": repr(o.min_frac), "hsc_compensate": str(o.compensate)})
        for l, lv in enumerate(hsc.levels()):
            p = f"hsc.{l:02d}."
            t[p + "F"] = lv["F"]
            t[p + "C"] = lv["C"]
            t[p + "dinv"] = lv["dinv"]
            t[p + "r_ptr"], t[p + "r_col"], t[p + "r_val"] = lv["R"]
            t[p + "p_ptr"], t[p + "p_col"], t[p + "p_val"] = lv["P"]
        Lc, inv = hsc.coarse()
        t["hsc.coarse_L"] = Lc
        t["hsc.coarse_inv"] = inv
    if influences is not None:
        start, idx, w = [np.asarray(a.cpu().numpy() if hasattr(a, "cpu") else a)
                         for a in influences]
        t["ddm.start"] = start.astype(np.int32)
        t["ddm.idx"] = idx.astype(np.int32)
        t["ddm.w"] = w
    write_safetensors(path, t, meta)


def _check_sidecar(st: SafeTensors, system):
    if st.metadata.get("format") != SIDECAR_FORMAT:
        raise ValueError("not a curvenet cache sidecar")
    if st.metadata.get("fingerprint") != system_fingerprint(system):
        raise ValueError("the sidecar was written for another system (A_ff differs)")


def load_hsc(path, system):
    """The hierarchy stored in a sidecar, rebuilt on the GPU without
    elimination or factorization (cn_hsc_import)."""
    from .curvenet import HSC
    st = read_safetensors(path)
    _check_sidecar(st, system)
    nl = int(st.metadata["hsc_levels"])
    lv_in = (L.HscLevelIn * max(1, nl))()
    keep = []
    for l in range(nl):
        p = f"hsc.{l:02d}."
        arrs = {}
        for k, dt in (("F", np.int32), ("C", np.int32), ("dinv", np.float64),
                      ("r_ptr", np.int32), ("r_col", np.int32), ("r_val", np.float64),
                      ("p_ptr", np.int32), ("p_col", np.int32), ("p_val", np.float64)):
            a = st.get(p + k)
            if a is None or a.dtype != np.dtype(dt):
                raise ValueError(f"sidecar: {p + k} missing or of the wrong dtype")
            arrs[k] = np.ascontiguousarray(a)
        keep.append(arrs)
        nF, nC = len(arrs["F"]), len(arrs["C"])
        if len(arrs["r_ptr"]) != nC + 1 or len(arrs["p_ptr"]) != nF + 1:
            raise ValueError(f"sidecar: level {l} pointer arrays of the wrong length")
        if (len(arrs["r_col"]) != arrs["r_ptr"][-1] or len(arrs["r_val"]) != arrs["r_ptr"][-1]
                or len(arrs["p_col"]) != arrs["p_ptr"][-1]
                or len(arrs["p_val"]) != arrs["p_ptr"][-1] or len(arrs["dinv"]) != nF):
            raise ValueError(f"sidecar: level {l} arrays of inconsistent lengths")
        x = lv_in[l]
        x.n, x.nF, x.nC = nF + nC, nF, nC
        for k in ("F", "C", "dinv", "r_ptr", "r_col", "r_val", "p_ptr", "p_col", "p_val"):
            setattr(x, k, arrs[k].ctypes.data if arrs[k].size else None)
    Lc, inv = st.get("hsc.coarse_L"), st.get("hsc.coarse_inv")
    if Lc is None or inv is None or Lc.ndim != 2:
        raise ValueError("sidecar: coarse factor missing or not square")
    nc = Lc.shape[0]
    if Lc.shape != (nc, nc) or inv.shape != (nc, nc):
        raise ValueError("sidecar: coarse factor missing or not square")
    # cn_hsc_import reads n_coarse^2 doubles from each
    if Lc.dtype != np.float64 or inv.dtype != np.float64:
        raise ValueError("sidecar: hsc.coarse_L / hsc.coarse_inv must be F64")
    Lc, inv = np.ascontiguousarray(Lc), np.ascontiguousarray(inv)
    o = L.HscOpts()
    L.load().cn_hsc_default_opts(C.byref(o))
    o.theta = float(st.metadata["hsc_theta"])
    o.coarse_max = int(st.metadata["hsc_coarse_max"])
    o.max_levels = int(st.metadata["hsc_max_levels"])
    o.min_frac = float(st.metadata["hsc_min_frac"])
    o.compensate = int(st.metadata["hsc_compensate"])
    return HSC.from_import(system, o, lv_in, nl, nc, Lc, inv)


def load_influences(path, system):
    """The skinning influences (start, idx, w) stored in a sidecar."""
    st = read_safetensors(path)
    _check_sidecar(st, system)
    start, idx, w = st.get("ddm.start"), st.get("ddm.idx"), st.get("ddm.w")
    if start is None or idx is None or w is None:
        raise ValueError("sidecar: ddm.start / ddm.idx / ddm.w missing")
    if start.dtype != np.int32 or idx.dtype != np.int32 or w.dtype not in (np.float64,
                                                                          np.float32):
        raise ValueError("sidecar: ddm.start and ddm.idx must be I32, ddm.w F64 or F32")
    start, idx, w = start.reshape(-1), idx.reshape(-1), w.reshape(-1)
    nv = system.info.nv
    check_influences(nv, start, idx, w)
    return start, idx, w


def check_influences(nv: int, start, idx, w):
    """The influence CSR must be consistent before a kernel walks it:
    start has nv + 1 entries, start[0] == 0, never decreases, and
    start[nv] == len(idx) == len(w)."""
    start = np.asarray(start).reshape(-1)
    if start.shape[0] != nv + 1:
        raise ValueError(f"influences: start has {start.shape[0]} entries, expected {nv + 1}")
    if start[0] != 0 or np.any(np.diff(start.astype(np.int64)) < 0):
        raise ValueError("influences: start must begin at 0 and never decrease")
    def count(a):
        return int(a.numel()) if hasattr(a, "numel") else int(np.asarray(a).size)
    if int(start[-1]) != count(idx) or int(start[-1]) != count(w):
        raise ValueError("influences: start[nv], len(idx) and len(w) differ")
