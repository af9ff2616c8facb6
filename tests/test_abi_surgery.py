"""CutAlgorithm's surgery KATs (CutAlgorithm.lean:259-507) run through the
library's cn_cut_* entry points.  These are host-side functions of the
C ABI (no device needed), so the test runs on the CPU; every result is
also compared array for array with the oracle's transcription."""
import ctypes

import numpy as np
import pytest

import oracle as O
from kat_data import QUAD, TRIANGLE, TWO_TRI_STRIP, he
from tool_godot_curvenet_amd import _lib

KEYS = ("target", "twin", "next", "face")


def _apply(mesh, ops):
    """Run ops [(name, args), ...] through the library on one halfedge
    mesh held in arrays with room for the new halfedges."""
    lib = _lib.load()
    h = he(mesh)
    extra = 2 * len(ops)
    arr = [np.concatenate([h[k], np.full(extra, -7, np.int32)]).astype(np.int32) for k in KEYS]
    nv, nf, nh = ctypes.c_int32(h["nv"]), ctypes.c_int32(h["nf"]), ctypes.c_int32(h["nh"])
    fns = {"subdivide": lib.cn_cut_subdivide_edge, "split": lib.cn_cut_split_face,
           "crack": lib.cn_cut_insert_crack}
    for name, args in ops:
        st = fns[name](ctypes.byref(nv), ctypes.byref(nf), ctypes.byref(nh), *args,
                       *(a.ctypes.data_as(ctypes.c_void_p) for a in arr))
        assert st == 0, (name, args, st)
    out = {k: a[:nh.value].copy() for k, a in zip(KEYS, arr)}
    out.update(nv=nv.value, nf=nf.value, nh=nh.value)
    return out


def _oracle(mesh, ops):
    m = he(mesh)
    for name, args in ops:
        m = {"subdivide": O.subdivide_edge, "split": O.split_face,
             "crack": O.insert_crack}[name](m, *args)
    return m


def loop_len(h, start):
    cur, steps = h["next"][start], 1
    while cur != start and steps < h["nh"]:
        cur = h["next"][cur]
        steps += 1
    return steps


def euler(h):
    return h["nv"] + (h["nf"] + 1) - h["nh"] // 2


CASES = {
    "triangleSplit": (TRIANGLE, [("subdivide", (0,))]),
    "quadSplit": (QUAD, [("subdivide", (0,))]),
    "triangleSplitTwice": (TRIANGLE, [("subdivide", (0,)), ("subdivide", (1,))]),
    "quadDiagonalSplit": (QUAD, [("split", (0, 2))]),
    "quadSubdivThenSplit": (QUAD, [("subdivide", (0,)), ("split", (1, 3))]),
    "triangleWithCrack": (TRIANGLE, [("crack", (0,))]),
    "quadWithCrack": (QUAD, [("crack", (0,))]),
    "twoTriStripWithChain": (TWO_TRI_STRIP, [("subdivide", (1,)), ("split", (2, 1)),
                                             ("split", (3, 4))]),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_library_equals_oracle(name):
    mesh, ops = CASES[name]
    got, want = _apply(mesh, ops), _oracle(mesh, ops)
    assert (got["nv"], got["nf"], got["nh"]) == (want["nv"], want["nf"], want["nh"])
    for k in KEYS:
        assert np.array_equal(got[k], want[k]), k
    assert O.he_manifold(got) == 15


def test_triangle_split():                              # CutAlgorithm.lean:269-312
    t = _apply(*CASES["triangleSplit"])
    assert (t["nv"], t["nh"], t["nf"]) == (4, 8, 1)
    assert loop_len(t, 0) == 4
    assert euler(he(TRIANGLE)) == euler(t)


def test_quad_split_and_twice():                        # CutAlgorithm.lean:314-328
    q = _apply(*CASES["quadSplit"])
    assert (q["nv"], q["nh"]) == (5, 10)
    t2 = _apply(*CASES["triangleSplitTwice"])
    assert (t2["nv"], t2["nh"]) == (5, 10)


def test_quad_diagonal_split():                         # CutAlgorithm.lean:330-375
    q = _apply(*CASES["quadDiagonalSplit"])
    assert (q["nv"], q["nh"], q["nf"]) == (4, 10, 2)
    assert loop_len(q, 1) == 3 and loop_len(q, 0) == 3
    assert euler(he(QUAD)) == euler(q)


def test_quad_subdiv_then_split():                      # CutAlgorithm.lean:377-385
    m = _apply(*CASES["quadSubdivThenSplit"])
    assert (m["nv"], m["nh"], m["nf"]) == (5, 12, 2)


def _partition_and_vtc(h, sample):
    """CutMesh.partitionOfUnity / VtCIsZero (CutMesh.lean:84-103) for a
    vertex-kind flag: every halfedge's target is exactly one of a mesh
    vertex (a V column) or a sample (a C column), so (V + C) 1 = 1 and
    V^T C = 0."""
    tgt = h["target"]
    V = np.zeros((h["nh"], h["nv"]))
    Cm = np.zeros((h["nh"], h["nv"]))
    for p, t in enumerate(tgt):
        (Cm if sample[t] else V)[p, t] = 1.0
    return bool(np.all(V.sum(1) + Cm.sum(1) == 1)), bool(np.all(V.T @ Cm == 0))


@pytest.mark.parametrize("name,nv,nh", [("quadSplit", 5, 10), ("quadWithCrack", 5, 10)])
def test_cutmesh_wrappers(name, nv, nh):                # CutAlgorithm.lean:387-416, 452-463
    """subdivideEdgeCM / insertCrackCM on the uncut quad with the new vertex
    tagged as a sample: counts, partition of unity and V^T C = 0."""
    m = _apply(*CASES[name])
    assert (m["nv"], m["nh"]) == (nv, nh)
    sample = np.zeros(m["nv"], bool)
    sample[m["nv"] - 1] = True
    assert _partition_and_vtc(m, sample) == (True, True)


def test_quad_split_cutmesh():                          # CutAlgorithm.lean:400-413
    m = _apply(*CASES["quadDiagonalSplit"])
    assert m["nh"] == 10
    assert _partition_and_vtc(m, np.zeros(m["nv"], bool)) == (True, True)
    # the two new halfedges are the appended pair 8, 9, each other's twin
    assert m["twin"][8] == 9 and m["twin"][9] == 8


def test_triangle_with_crack():                         # CutAlgorithm.lean:419-450
    t = _apply(*CASES["triangleWithCrack"])
    assert (t["nv"], t["nh"], t["nf"]) == (4, 8, 1)
    assert t["face"][6] == 0 and t["face"][7] == 0
    assert t["twin"][6] == 7 and t["twin"][7] == 6
    assert loop_len(t, 0) == 5


def test_chain_across_two_faces():                      # CutAlgorithm.lean:465-507
    m = _apply(*CASES["twoTriStripWithChain"])
    assert (m["nv"], m["nh"], m["nf"]) == (5, 16, 4)
    assert loop_len(m, 0) == 3
    # every face of the chain is a triangle
    for f in range(m["nf"]):
        first = int(np.flatnonzero(m["face"] == f)[0])
        assert loop_len(m, first) == 3


def test_bad_halfedge_rejected():
    lib = _lib.load()
    h = he(TRIANGLE)
    arr = [np.concatenate([h[k], np.zeros(2, np.int32)]).astype(np.int32) for k in KEYS]
    nv, nf, nh = ctypes.c_int32(3), ctypes.c_int32(1), ctypes.c_int32(6)
    st = lib.cn_cut_subdivide_edge(ctypes.byref(nv), ctypes.byref(nf), ctypes.byref(nh), 6,
                                   *(a.ctypes.data_as(ctypes.c_void_p) for a in arr))
    assert st == _lib.CN_BAD_ARG and (nv.value, nh.value) == (3, 6)
