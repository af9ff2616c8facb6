"""The C-ABI library loads without a GPU and exports every symbol that
include/curvenet_la.h declares (no compute calls here)."""
import ctypes

from tool_godot_curvenet_amd import _lib


def test_library_exports_declared_symbols():
    lib = _lib.load()
    names = _lib.declared_symbols()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_bindings_cover_header():
    assert set(_lib.declared_symbols()) == set(_lib._SIGS)


def test_version_and_opts():
    lib = _lib.load()
    assert lib.cn_version() >= 100
    o = _lib.PcgOpts()
    lib.cn_pcg_default_opts(ctypes.byref(o))
    assert o.tol_rel == 1e-9 and o.precision == _lib.CN_PREC_MIXED


def test_host_surgery_matches_spec():
    """cn_cut_* are host-side (no device needed): subdivideEdge on the spec's
    triangle (CutAlgorithm.lean:270-277)."""
    import numpy as np

    import oracle as O
    from kat_data import TRIANGLE, he
    lib = _lib.load()
    h = he(TRIANGLE)
    arr = [np.concatenate([h[k], np.zeros(2, np.int32)]).astype(np.int32)
           for k in ("target", "twin", "next", "face")]
    nv, nf, nh = ctypes.c_int32(3), ctypes.c_int32(1), ctypes.c_int32(6)
    st = lib.cn_cut_subdivide_edge(ctypes.byref(nv), ctypes.byref(nf), ctypes.byref(nh), 0,
                                   *(a.ctypes.data_as(ctypes.c_void_p) for a in arr))
    assert st == 0 and (nv.value, nh.value) == (4, 8)
    o = O.subdivide_edge(h, 0)
    for a, k in zip(arr, ("target", "twin", "next", "face")):
        assert np.array_equal(a, o[k])
    for fn, args, okey in ((lib.cn_cut_split_face, (0, 2), O.split_face(he(TRIANGLE), 0, 2)),
                           (lib.cn_cut_insert_crack, (0,), O.insert_crack(he(TRIANGLE), 0))):
        h = he(TRIANGLE)
        arr = [np.concatenate([h[k], np.zeros(2, np.int32)]).astype(np.int32)
               for k in ("target", "twin", "next", "face")]
        nv, nf, nh = ctypes.c_int32(3), ctypes.c_int32(1), ctypes.c_int32(6)
        assert fn(ctypes.byref(nv), ctypes.byref(nf), ctypes.byref(nh), *args,
                  *(a.ctypes.data_as(ctypes.c_void_p) for a in arr)) == 0
        for a, k in zip(arr, ("target", "twin", "next", "face")):
            assert np.array_equal(a, okey[k]), k
