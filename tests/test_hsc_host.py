"""The host builder of the hierarchical Schur preconditioner's levels
(tool_godot_curvenet_amd/csrc/cn_hsc_host.cpp) on the CPU: every level equal
bit for bit to the oracle's hsc_build restatement (F, C, 1/a_ff, R, P, and
the coarse matrix the levels leave), with one thread and with several (the
row-parallel passes must not change any sum's order)."""
import ctypes as C
import pathlib
import subprocess

import numpy as np
import pytest

import oracle as O
from tool_godot_curvenet_amd import meshgen as G

ROOT = pathlib.Path(__file__).resolve().parents[1]
CSRC = ROOT / "tool_godot_curvenet_amd" / "csrc"
SRC = [ROOT / "tests" / "native" / "hsc_host_check.cpp", CSRC / "cn_hsc_host.cpp"]
OUT = ROOT / "tests" / "_build" / "libhsc_host_check.so"
FIELDS = (("F", np.int32), ("C", np.int32), ("dinv", np.float64), ("r_ptr", np.int32),
          ("r_col", np.int32), ("r_val", np.float64), ("p_ptr", np.int32),
          ("p_col", np.int32), ("p_val", np.float64), ("next_ptr", np.int32),
          ("next_col", np.int32), ("next_val", np.float64))


@pytest.fixture(scope="module")
def host():
    OUT.parent.mkdir(exist_ok=True)
    deps = SRC + [CSRC / "cn_hsc_host.h", ROOT / "include" / "curvenet_la.h"]
    if not OUT.exists() or any(d.stat().st_mtime > OUT.stat().st_mtime for d in deps):
        subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-pthread", "-fPIC",
                        "-shared", f"-I{ROOT / 'include'}", f"-I{CSRC}", *map(str, SRC),
                        "-o", str(OUT)], check=True)
    lib = C.CDLL(str(OUT))
    vp = C.c_void_p
    lib.hsc_level_host.restype = vp
    lib.hsc_level_host.argtypes = [C.c_int32, vp, vp, vp, C.c_double, C.c_int32, C.c_double,
                                   C.c_int32, C.POINTER(C.c_int32)]
    lib.hsc_level_size.restype = C.c_int64
    lib.hsc_level_size.argtypes = [vp, C.c_int32]
    lib.hsc_level_copy.argtypes = [vp, C.c_int32, vp]
    lib.hsc_level_free.argtypes = [vp]
    return lib


def host_level(lib, ptr, col, val, theta, comp, threads, min_frac=0.05):
    ptr = np.ascontiguousarray(ptr, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    val = np.ascontiguousarray(val, np.float64)
    err = C.c_int32(0)
    h = lib.hsc_level_host(len(ptr) - 1, ptr.ctypes.data, col.ctypes.data, val.ctypes.data,
                           theta, 1 if comp == "diag" else 0, min_frac, threads, C.byref(err))
    assert err.value == 0
    if not h:
        return None
    out = {}
    for i, (name, dt) in enumerate(FIELDS):
        a = np.empty(lib.hsc_level_size(h, i), dt)
        if a.size:
            lib.hsc_level_copy(h, i, a.ctypes.data)
        out[name] = a
    lib.hsc_level_free(h)
    return out


def _matrix(name):
    m = G.config_mesh(name)
    nt, origin, con, _ = O.cut_edges(m.nv, m.tris, m.cut_edges)
    pos = m.pos[origin]
    h = O.he_from_triangles(pos.shape[0], nt)
    return O.cut_laplacian_csr(h, con, pos)["A_ff"]


def _levels(lib, A, theta, comp, threads, coarse_max):
    ptr, col, val = A
    out = []
    while len(ptr) - 1 > coarse_max and len(out) < 24:
        L = host_level(lib, ptr, col, val, theta, comp, threads)
        if L is None:
            break
        out.append(L)
        ptr, col, val = L["next_ptr"], L["next_col"], L["next_val"]
    return out, (ptr, col, val)


@pytest.mark.parametrize("name,theta,comp", [("cylinder64", 0.015, "lump"),
                                             ("cylinder64", 0.05, "diag"),
                                             ("sphere45", 0.015, "lump")])
def test_levels_equal_oracle(host, name, theta, comp):
    A = _matrix(name)
    ref = O.hsc_build(*A, theta=theta, coarse_max=256, comp=comp)
    for threads in (1, 6):
        got, (cp, cc, cv) = _levels(host, A, theta, comp, threads, coarse_max=256)
        assert len(got) == len(ref["levels"])
        for L, R in zip(got, ref["levels"]):
            assert np.array_equal(L["F"], R["F"]) and np.array_equal(L["C"], R["C"])
            assert np.array_equal(L["dinv"], R["dinv"])
            rp, rc, rv = R["R"]
            # the oracle numbers R's columns within F, the builder by level row
            assert np.array_equal(L["r_ptr"], rp) and np.array_equal(L["r_col"], R["F"][rc])
            assert np.array_equal(L["r_val"], rv)
            pp, pc, pv = R["P"]
            assert np.array_equal(L["p_ptr"], pp) and np.array_equal(L["p_col"], pc)
            assert np.array_equal(L["p_val"], pv)
        n = len(cp) - 1
        D = np.zeros((n, n))
        for i in range(n):
            D[i, cc[cp[i]:cp[i + 1]]] = cv[cp[i]:cp[i + 1]]
        assert np.array_equal(D, ref["coarse"])


def test_thread_counts_agree_on_the_capsule_union(host):
    """The 100k-row capsule union: 1, 3 and 16 threads give identical levels
    (many row chunks per pass)."""
    A = _matrix("capsules100k")
    base, _ = _levels(host, A, 0.015, "lump", 1, coarse_max=512)
    for threads in (3, 16):
        got, _ = _levels(host, A, 0.015, "lump", threads, coarse_max=512)
        assert len(got) == len(base)
        for L, B in zip(got, base):
            for k, _ in FIELDS:
                assert np.array_equal(L[k], B[k]), k


def test_planted_reassociation_fails(host):
    """Planted mistake in the oracle's rule: the Schur term a_cf (a_fc2 / a_ff)
    instead of (a_cf a_fc2) / a_ff changes bits the builder must not have."""
    A = _matrix("cylinder64")
    got, _ = _levels(host, A, 0.015, "lump", 4, coarse_max=256)
    ptr, col, val = (np.asarray(a) for a in A)
    ref = O.hsc_build(ptr, col, val, theta=0.015, coarse_max=256)
    n = len(ptr) - 1
    rows = [dict(zip(col[ptr[i]:ptr[i + 1]].tolist(), val[ptr[i]:ptr[i + 1]].tolist()))
            for i in range(n)]
    F = ref["levels"][0]["F"].tolist()
    C_ = ref["levels"][0]["C"].tolist()
    isF = np.zeros(n, bool)
    isF[F] = True
    cidx = {c: t for t, c in enumerate(C_)}
    S = [{cidx[j]: a for j, a in rows[c].items() if not isF[j]} for c in C_]
    for f in F:
        d = rows[f][f]
        nb = sorted(j for j in rows[f] if j != f)
        for c in nb:
            for c2 in nb:
                t = cidx[c2]
                S[cidx[c]][t] = S[cidx[c]].get(t, 0.0) - rows[c][f] * (rows[f][c2] / d)
    nxt = got[0]
    diffs = same = 0
    for t in range(len(C_)):
        for q in range(nxt["next_ptr"][t], nxt["next_ptr"][t + 1]):
            j = nxt["next_col"][q]
            if j != t:
                if nxt["next_val"][q] != S[t][j]:
                    diffs += 1
                else:
                    same += 1
    assert diffs > 0 and same > 0
