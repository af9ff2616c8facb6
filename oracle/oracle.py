"""ctypes front end of liboracle.so (cn_oracle.c), numpy in / numpy out.

TEST INFRASTRUCTURE ONLY: the checker the GPU path is compared against.
Every function names the spec definition it restates (see cn_oracle.c for
file:line citations).  Parity status of each function is listed in
PINS below: every function is pinned by at least one independent check in
tests/test_oracle_pins.py or tests/test_oracle_kat.py.
"""
from __future__ import annotations

import ctypes as C
import math
import pathlib
import subprocess

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"

# function -> the independent checks that pin it
PINS = {
    "he_from_triangles_spec": "spec KATs (counts, manifold?) HalfedgeBuilder.lean:115-136",
    "he_from_triangles": "index-identical to the literal spec builder on random meshes",
    "he_manifold": "spec KATs + planted defects",
    "subdivide_edge/split_face/insert_crack": "CutAlgorithm.lean:269-507 KATs",
    "cut_edges": "cylinder seam closed form (64 copies at nv+j), Euler/partition invariants, "
                 "brute-force wedge recount",
    "cot_at_vertex/triangle_cot_laplacian/polygon_cot_laplacian": "closed forms (equilateral -1/(2 sqrt 3)), row sums, symmetry",
    "mollify_triangle/cot_from_lengths/robust_triangle_cot_laplacian": "RobustLaplacian.lean:160-290 KATs (cot 60/30/90 degrees, "
                                 "mollification margins, equilateral == embedded, finite on a "
                                 "collinear triangle); equal to the embedded cot on random "
                                 "non-degenerate triangles with delta = 0 (law of cosines + "
                                 "Heron vs cross product)",
    "assemble_vtlhv_dense/harmonic_rhs_dense": "CutMeshLaplacian.lean:101-179 KATs, linear precision",
    "cut_laplacian_csr": "bit-identical to the dense spec transcription on small cut meshes",
    "harmonic_rhs_csr": "bit-identical to the dense spec transcription; HarmonicSolve KATs",
    "solve_deformation_dense": "DeformSolve.lean:164-253 KATs (translation, rotation, scale on "
                               "triangleWithSample); EndToEndExample Bezier-driven translation",
    "deform_rhs_csr/deform_rhs_dense": "bit-identical to the dense spec transcription (deform_rhs_dense); a "
                      "rigid motion is reproduced exactly by the two-stage solve",
    "bezier3/isolated_segment_gradient/deformation_gradient/mat3_inv": "ScaledFrames.lean and "
        "EndToEndExample.lean KATs (identity, rotation, scale, rotate+scale, det, inverse)",
    "spmv/spmm/dot/csr_diag": "identity KAT; numpy reference",
    "cg": "SparseLinAlg.lean:171-287 KATs; dense GE",
    "ic0_factor/ic0_factor3/permute_csr": "LL^T == A on the kept pattern; IncompleteCholesky.lean:128-186 KATs; "
                  "full-pattern IC(0) == dense Cholesky",
    "ic0_solve": "LL^T x == b",
    "pcg": "dense GE / numpy solve",
    "pcg_many": "column for column equal to pcg (the same arithmetic on each thread)",
    "greedy_color/color_classes": "GraphColoring.lean:79-117 KATs, validity",
    "sgs_sweep/sgs_apply": "equal to lexical Gauss-Seidel in colour order (the kernel's stated "
                           "equivalence); one sweep from 0 == the textbook SGS preconditioner "
                           "(D+L) D^-1 (D+U) solved densely; many sweeps converge to the "
                           "dense solve; zero-diagonal rows untouched; planted mistakes fail",
    "restrict_sum": "numpy np.add.at closed form (exact for the same order), hand KAT, "
                    "parents outside [0, n_coarse) ignored",
    "hsc_build/hsc_apply": "theta = 0: one level == dense Schur complement, exact hierarchy "
                           "apply == A^-1; lump/diag compensation invariants; planted sign",
    "dense_solve/dense_solve_multi/lu_factor/lu_solve/cholesky_factor/cholesky_solve": "DenseLinAlg.lean:294-489 KATs, numpy.linalg",
    "hsc_eliminate_vertex/hsc_edge_weight/hsc_degree_weight": "HSC KATs (HierarchicalSparsifyCompensate.lean:134-231), "
                            "dense Schur complement",
    "apply_transform4x4/lbs_matvec": "DirectDeltaMush.lean:115-165 and "
        "Features/KnownKnowns/DirectDeltaMush.lean KATs (identity, translation, blends, empty "
        "influences); rotation about an axis; linear in the weights; planted mistakes fail",
    "smooth_weights": "DirectDeltaMushBind.lean:169-197 KATs (nu = 0 identity, half weights "
                      "fixed, partition of unity); dense matrix form "
                      "((1-omega) I + omega D^-1 A)^nu on random graphs; planted mistakes fail",
    "sparsify_topk": "DirectDeltaMushBind.lean:199-262 KATs (top-2 of [0.4 .. 0.1] = 4/7, 3/7; "
                     "K = width; smooth-then-sparsify partition of unity); a literal Python "
                     "selection sort with swaps on tie-heavy rows; top-K set == numpy "
                     "argsort on distinct values",
    "project_to_vertices/projected_positions": "SurfaceProjection.lean:99-130 KATs (knot at "
        "a vertex, nearest of three, all kinds vertex, empty knots, empty mesh -> -1 and the "
        "knot's own position); numpy first-argmin of the squared distances on random clouds "
        "with planted duplicate vertices and equidistant knots (lowest index wins); a planted "
        "<= (last index wins) fails",
    "promote_vertex_samples": "SurfaceProjection.lean:132-157 KATs (shared column, encounter "
        "order, order matters, repeated knot); closed form on random projections: columns "
        "are the ranks of np.unique(return_index) first occurrences; planted 'column = "
        "vertex order' fails",
    "adjoint": "central finite differences (h = 1e-6) of a quadratic loss through a dense "
               "FP64 solve, in x_c and in every edge weight; planted sign / transpose "
               "mistakes fail them",
}


def build() -> pathlib.Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


_lib = None
i32, i64, f64 = C.c_int32, C.c_int64, C.c_double
P = C.c_void_p


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        _lib = C.CDLL(str(LIB))
        _lib.orc_cot_at_vertex.restype = f64
        _lib.orc_dot.restype = f64
        _lib.orc_hsc_degree_weight.restype = f64
        _lib.orc_hsc_edge_weight.restype = f64
        _lib.orc_pcg_many.restype = i64
        _lib.orc_mat3_det.restype = f64
        _lib.orc_mollify_triangle.restype = f64
        _lib.orc_mollify_triangle.argtypes = [f64, f64, f64, f64]
        _lib.orc_cot_from_lengths.restype = f64
        _lib.orc_cot_from_lengths.argtypes = [f64, f64, f64]
        _lib.orc_set_laplacian.argtypes = [i32, f64]
        _lib.orc_robust_triangle_cot_laplacian.argtypes = [f64, P, P, P, P]
        for name in ("orc_he_from_triangles_spec", "orc_he_from_triangles", "orc_he_manifold",
                     "orc_subdivide_edge", "orc_split_face", "orc_insert_crack", "orc_cut_edges",
                     "orc_cut_laplacian_csr", "orc_cg", "orc_ic0_factor", "orc_pcg",
                     "orc_greedy_color", "orc_hsc_eliminate_vertex"):
            getattr(_lib, name).restype = i32
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(P)


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ------------------------------------------------------------------ topology --
def he_from_triangles(nv, tris, spec=False):
    tris = _i(np.asarray(tris).reshape(-1, 3))
    nf = tris.shape[0]
    cap = max(6 * nf, 1)
    t, w, n, f = (np.zeros(cap, np.int32) for _ in range(4))
    fn = lib().orc_he_from_triangles_spec if spec else lib().orc_he_from_triangles
    nh = fn(nv, nf, _p(tris), _p(t), _p(w), _p(n), _p(f))
    return {"target": t[:nh].copy(), "twin": w[:nh].copy(), "next": n[:nh].copy(),
            "face": f[:nh].copy(), "nv": nv, "nf": nf, "nh": nh}


def he_manifold(he) -> int:
    t, w, n, f = (_i(he[k]) for k in ("target", "twin", "next", "face"))
    return lib().orc_he_manifold(he["nv"], he["nf"], he["nh"], _p(t), _p(w), _p(n), _p(f))


def _surgery(fn, he, *args):
    nh = he["nh"]
    arrs = [np.concatenate([_i(he[k]), np.zeros(2, np.int32)])
            for k in ("target", "twin", "next", "face")]
    nv, nf = C.c_int32(he["nv"]), C.c_int32(he["nf"])
    nh2 = fn(C.byref(nv), C.byref(nf), nh, *args, *(_p(a) for a in arrs))
    return {"target": arrs[0][:nh2], "twin": arrs[1][:nh2], "next": arrs[2][:nh2],
            "face": arrs[3][:nh2], "nv": nv.value, "nf": nf.value, "nh": nh2}


def subdivide_edge(he, h):
    return _surgery(lib().orc_subdivide_edge, he, h)


def split_face(he, a, b):
    return _surgery(lib().orc_split_face, he, a, b)


def insert_crack(he, h):
    return _surgery(lib().orc_insert_crack, he, h)


def cut_edges(nv, tris, edges):
    tris = _i(np.asarray(tris).reshape(-1, 3))
    edges = _i(np.asarray(edges).reshape(-1, 2))
    cap = nv + 2 * edges.shape[0] + 1
    nt = np.zeros_like(tris)
    origin = np.zeros(cap, np.int32)
    con = np.zeros(cap, np.uint8)
    nv2 = lib().orc_cut_edges(nv, tris.shape[0], _p(tris), edges.shape[0], _p(edges), _p(nt),
                              _p(origin), _p(con))
    if nv2 < 0:
        raise ValueError(f"orc_cut_edges failed ({nv2})")
    return nt, origin[:nv2].copy(), con[:nv2].copy(), nv2


def project_to_vertices(knots, pos):
    """SurfaceProjection.projectToVertices (SurfaceProjection.lean:46-64),
    direct transcription (strict <, lowest index wins)."""
    knots = _d(np.asarray(knots).reshape(-1, 3))
    pos = _d(np.asarray(pos).reshape(-1, 3))
    out = np.empty(knots.shape[0], np.int32)
    for i, k in enumerate(knots):
        best, bi = 1e308, -1
        d = pos - k[None]
        dd = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]).tolist()
        for v in range(pos.shape[0]):
            if dd[v] < best:
                best, bi = dd[v], v
        out[i] = bi
    return out


def projected_positions(knots, pos, index):
    """The `position` field of SurfaceProjection.projectToVertices
    (SurfaceProjection.lean:58): the chosen mesh vertex, or the knot itself
    when no vertex was found (empty mesh, index -1)."""
    knots = _d(np.asarray(knots).reshape(-1, 3))
    pos = _d(np.asarray(pos).reshape(-1, 3))
    index = np.asarray(index, np.int64)
    out = knots.copy()
    ok = index >= 0
    out[ok] = pos[index[ok]]
    return out


def promote_vertex_samples(vertex_count, index, kinds=None):
    """SurfaceProjection.promoteVertexSamples (SurfaceProjection.lean:64-80),
    direct transcription: walking the projections in order, the first knot
    that lands on a vertex gives that vertex the next column; later knots on
    the same vertex share it; knots with no vertex (meshIndex < 0) or of a
    non-vertex kind keep -1.  `kinds` (0 = vertex) defaults to all vertex,
    the only kind projectToVertices produces.  Returns (knot_to_col,
    vertex_to_col, ncols).  An index >= vertex_count is rejected (the spec's
    `vertexToCol[v]!` would panic there)."""
    index = [int(v) for v in np.asarray(index).reshape(-1)]
    kinds = [0] * len(index) if kinds is None else [int(t) for t in kinds]
    vertex_to_col = [-1] * int(vertex_count)
    knot_to_col = [-1] * len(index)
    next_col = 0
    for i, v in enumerate(index):
        if kinds[i] == 0 and v >= 0:
            if v >= vertex_count:
                raise IndexError(f"knot {i} projects to vertex {v} >= {vertex_count}")
            cur = vertex_to_col[v]
            if cur < 0:
                vertex_to_col[v] = next_col
                knot_to_col[i] = next_col
                next_col += 1
            else:
                knot_to_col[i] = cur
    return (np.asarray(knot_to_col, np.int32), np.asarray(vertex_to_col, np.int32), next_col)


# ----------------------------------------------------------------- Laplacian --
def cot_at_vertex(apex, o1, o2):
    return lib().orc_cot_at_vertex(_p(_d(apex)), _p(_d(o1)), _p(_d(o2)))


def triangle_cot_laplacian(a, b, c):
    m = np.zeros(9)
    lib().orc_triangle_cot_laplacian(_p(_d(a)), _p(_d(b)), _p(_d(c)), _p(m))
    return m.reshape(3, 3)


def polygon_cot_laplacian(poly):
    poly = _d(np.asarray(poly).reshape(-1, 3))
    n = poly.shape[0]
    m = np.zeros(n * n)
    lib().orc_polygon_cot_laplacian(n, _p(poly), _p(m))
    return m.reshape(n, n)


def assemble_vtlhv_dense(he, is_sample, pos):
    nv = he["nv"]
    out = np.zeros(nv * nv)
    lib().orc_assemble_vtlhv_dense(nv, he["nf"], he["nh"], _p(_i(he["target"])),
                                   _p(_i(he["next"])), _p(_i(he["face"])),
                                   _p(np.ascontiguousarray(is_sample, np.uint8)),
                                   _p(_d(pos)), _p(out))
    return out.reshape(nv, nv)


def harmonic_rhs_dense(he, col_of_vertex, pos, Fc):
    Fc = _d(np.asarray(Fc))
    k = Fc.shape[1] if Fc.ndim == 2 else 1
    nv = he["nv"]
    out = np.zeros(nv * k)
    lib().orc_harmonic_rhs_dense(nv, he["nf"], he["nh"], _p(_i(he["target"])), _p(_i(he["next"])),
                                 _p(_i(he["face"])), _p(_i(col_of_vertex)), _p(_d(pos)), k,
                                 _p(Fc), _p(out))
    return out.reshape(nv, k)


def mollify_triangle(delta, eA, eB, eC):
    """RobustLaplacian.mollifyTriangle (RobustLaplacian.lean:83-89)."""
    return lib().orc_mollify_triangle(delta, eA, eB, eC)


def cot_from_lengths(eo, e1, e2):
    """RobustLaplacian.cotFromLengths (RobustLaplacian.lean:106-109)."""
    return lib().orc_cot_from_lengths(eo, e1, e2)


def robust_triangle_cot_laplacian(delta, a, b, c):
    """RobustLaplacian.robustTriangleCotLaplacian (RobustLaplacian.lean:136-140)."""
    m = np.zeros(9)
    lib().orc_robust_triangle_cot_laplacian(delta, _p(_d(a)), _p(_d(b)), _p(_d(c)), _p(m))
    return m.reshape(3, 3)


class laplacian:
    """Context: the oracle assembles with the robust triangle Laplacian
    (margin delta) inside the block, the embedded cot outside."""

    def __init__(self, kind="robust", delta=0.0):
        self.kind, self.delta = kind, delta

    def __enter__(self):
        lib().orc_set_laplacian(1 if self.kind == "robust" else 0, self.delta)
        return self

    def __exit__(self, *exc):
        lib().orc_set_laplacian(0, 0.0)
        return False


def cut_laplacian_csr(he, constrained, pos):
    """Returns dict with free_of, con_of, A_ff=(ptr,col,val), A_fc=(ptr,col,val)."""
    nv, nh = he["nv"], he["nh"]
    con = np.ascontiguousarray(constrained, np.uint8)
    cap = nv + 6 * he["nf"] + 16
    free_of = np.zeros(nv, np.int32)
    con_of = np.zeros(nv, np.int32)
    ffp, fcp = np.zeros(nv + 1, np.int32), np.zeros(nv + 1, np.int32)
    ffc, fcc = np.zeros(cap, np.int32), np.zeros(cap, np.int32)
    ffv, fcv = np.zeros(cap), np.zeros(cap)
    nnz = np.zeros(2, np.int64)
    nfree = lib().orc_cut_laplacian_csr(nv, he["nf"], nh, _p(_i(he["target"])), _p(_i(he["next"])),
                                        _p(_i(he["face"])), _p(con), _p(_d(pos)), _p(free_of),
                                        _p(con_of), _p(ffp), _p(ffc), _p(ffv), _p(fcp), _p(fcc),
                                        _p(fcv), C.c_int64(cap), _p(nnz))
    if nfree < 0:
        raise ValueError("orc_cut_laplacian_csr overflow")
    a, b = int(nnz[0]), int(nnz[1])
    return {"n_free": nfree, "free_of": free_of, "con_of": con_of,
            "A_ff": (ffp[:nfree + 1].copy(), ffc[:a].copy(), ffv[:a].copy()),
            "A_fc": (fcp[:nfree + 1].copy(), fcc[:b].copy(), fcv[:b].copy())}


def apply_transform4x4(T, x, y, z):
    """DirectDeltaMush.applyTransform4x4 (DirectDeltaMush.lean:69-73); T is 16
    row-major values."""
    out = np.zeros(3)
    lib().orc_apply_transform4x4(_p(_d(T)), C.c_double(x), C.c_double(y), C.c_double(z),
                                 _p(out))
    return out


def lbs_matvec(rest, T, start, idx, w):
    """DirectDeltaMush.lbsMatvec (DirectDeltaMush.lean:85-98) for every vertex:
    rest nv x 3, T nt x 16, influences CSR (start, idx, w)."""
    rest = _d(rest).reshape(-1, 3)
    T = _d(T).reshape(-1, 16)
    out = np.zeros_like(rest)
    lib().orc_lbs_matvec(len(rest), _p(rest), _p(T), len(T), _p(_i(start)), _p(_i(idx)),
                         _p(_d(w)), _p(out))
    return out


def smooth_weights(W, adj_ptr, adj, nu, omega):
    """DirectDeltaMushBind.smoothWeights (DirectDeltaMushBind.lean:76-97);
    adjacency as CSR (adj_ptr, adj), neighbour order kept."""
    W = _d(W)
    n, nh = W.shape
    out = np.zeros_like(W)
    lib().orc_smooth_weights(n, nh, _p(_i(adj_ptr)), _p(_i(adj)), _p(W), int(nu),
                             C.c_double(omega), _p(out))
    return out


def sparsify_topk(W, K):
    """DirectDeltaMushBind.sparsifyTopK (DirectDeltaMushBind.lean:107-136):
    (idx, w), each n x min(K, nh)."""
    W = _d(W)
    n, nh = W.shape
    take = min(K, nh)
    idx = np.zeros((n, take), np.int32)
    w = np.zeros((n, take))
    lib().orc_sparsify_topk(n, nh, _p(W), int(K), _p(idx), _p(w))
    return idx, w


def bezier3(P, t):
    """EndToEndExample.bezier3 (EndToEndExample.lean:52-60); P is 4 x 3."""
    out = np.zeros(3)
    lib().orc_bezier3(_p(_d(P)), C.c_double(t), _p(out))
    return out


def isolated_segment_gradient(rest_p, rest_q, posed_p, posed_q):
    """ScaledFrames.isolatedSegmentGradient (ScaledFrames.lean:90-95), 3 x 3."""
    F = np.zeros(9)
    lib().orc_isolated_segment_gradient(_p(_d(rest_p)), _p(_d(rest_q)), _p(_d(posed_p)),
                                        _p(_d(posed_q)), _p(F))
    return F.reshape(3, 3)


def deformation_gradient(rest, posed):
    """ScaledFrames.deformationGradient (ScaledFrames.lean:147-150)."""
    F = np.zeros(9)
    lib().orc_deformation_gradient(_p(_d(rest)), _p(_d(posed)), _p(F))
    return F.reshape(3, 3)


def mat3_inv(m):
    o = np.zeros(9)
    lib().orc_mat3_inv(_p(_d(m)), _p(o))
    return o.reshape(3, 3)


def deform_rhs_dense(he, col_of_vertex, pos, F, Xc):
    """Stage-2 RHS -V^T Lh (C Xc - y_h) (DeformSolve.lean:118-133), dense
    spec form; F per vertex (nv x 9P, samples hold F_c), Xc nc x 3P."""
    F, Xc = _d(np.asarray(F)), _d(np.asarray(Xc))
    P = F.shape[1] // 9
    out = np.zeros(he["nv"] * 3 * P)
    lib().orc_deform_rhs_dense(he["nv"], he["nf"], he["nh"], _p(_i(he["target"])),
                               _p(_i(he["next"])), _p(_i(he["face"])), _p(_i(col_of_vertex)),
                               _p(_d(pos)), P, _p(F), _p(Xc), _p(out))
    return out.reshape(he["nv"], 3 * P)


def deform_rhs_csr(he, free_of, con_of, pos, Ffree, Fc, Xc):
    """The stage-2 RHS in the sparse assembly order, free rows only."""
    Ffree, Fc, Xc = _d(np.asarray(Ffree)), _d(np.asarray(Fc)), _d(np.asarray(Xc))
    P = Ffree.shape[1] // 9
    nfree = int((np.asarray(free_of) >= 0).sum())
    out = np.zeros(nfree * 3 * P)
    lib().orc_deform_rhs_csr(he["nv"], he["nh"], _p(_i(he["target"])), _p(_i(he["next"])),
                             _p(_i(he["face"])), _p(_i(free_of)), _p(_i(con_of)), _p(_d(pos)),
                             P, _p(Ffree), _p(Fc), _p(Xc), _p(out))
    return out.reshape(nfree, 3 * P)


def solve_deformation_dense(he, is_sample, pos, Fc, Xc):
    """DeformSolve.solveDeformation (DeformSolve.lean:110-135), the dense
    transcription: stage 1 (k = 9), computeFh, averageOverFaces, computeYh,
    stage 2 (k = 3), each solve by DenseLinAlg.solveMulti.  One pose; the
    sample rows of the result are 0 (the caller overlays them)."""
    is_sample = np.asarray(is_sample, np.uint8)
    col = np.where(is_sample > 0, np.cumsum(is_sample) - 1, -1).astype(np.int32)
    lhs = assemble_vtlhv_dense(he, is_sample, pos)
    Fv = dense_solve_multi(lhs, harmonic_rhs_dense(he, col, pos, Fc))
    F = Fv.copy()
    F[is_sample > 0] = np.asarray(Fc).reshape(-1, 9)      # computeFh: samples take F_c
    return dense_solve_multi(lhs, deform_rhs_dense(he, col, pos, F, Xc))


def harmonic_rhs_csr(he, free_of, con_of, pos, Xc):
    Xc = _d(np.asarray(Xc))
    k = Xc.shape[1] if Xc.ndim == 2 else 1
    nfree = int((np.asarray(free_of) >= 0).sum())
    out = np.zeros(nfree * k)
    lib().orc_harmonic_rhs_csr(he["nv"], he["nh"], _p(_i(he["target"])), _p(_i(he["next"])),
                               _p(_i(he["face"])), _p(_i(free_of)), _p(_i(con_of)), _p(_d(pos)),
                               k, _p(Xc), _p(out))
    return out.reshape(nfree, k)


# ------------------------------------------------------------ sparse linalg --
def spmv(ptr, col, val, x):
    n = len(ptr) - 1
    y = np.zeros(n)
    lib().orc_spmv(n, _p(_i(ptr)), _p(_i(col)), _p(_d(val)), _p(_d(x)), _p(y))
    return y


def spmm(ptr, col, val, X):
    X = _d(X)
    n, k = len(ptr) - 1, X.shape[1]
    Y = np.zeros(n * k)
    lib().orc_spmm(n, _p(_i(ptr)), _p(_i(col)), _p(_d(val)), k, _p(X), _p(Y))
    return Y.reshape(n, k)


def dot(a, b):
    a, b = _d(a), _d(b)
    return lib().orc_dot(len(a), _p(a), _p(b))


def cg(ptr, col, val, b, max_iter, tol, x0=None):
    """SparseLinAlg.cg / cgWithGuess. Returns (x, iterations)."""
    n = len(ptr) - 1
    x = np.zeros(n)
    b = _d(b)
    it = lib().orc_cg(n, _p(_i(ptr)), _p(_i(col)), _p(_d(val)), _p(b),
                      _p(_d(x0)) if x0 is not None else None, max_iter, C.c_double(tol), _p(x))
    return x, it


def ic0_factor3(A, pattern):
    L = np.zeros(9)
    lib().orc_ic0_factor3(_p(_d(np.asarray(A).reshape(-1))), _p(_d(np.asarray(pattern).reshape(-1))),
                          _p(L))
    return L.reshape(3, 3)


def ic0_factor(ptr, col, val):
    """CSR IC(0); returns (l_ptr, l_col, l_val) incl. diagonal (last per row)."""
    n = len(ptr) - 1
    ptr, col, val = _i(ptr), _i(col), _d(val)
    nl = int(sum(1 for i in range(n) for p in range(ptr[i], ptr[i + 1]) if col[p] <= i))
    lp = np.zeros(n + 1, np.int32)
    lc = np.zeros(max(nl, 1), np.int32)
    lv = np.zeros(max(nl, 1))
    st = lib().orc_ic0_factor(n, _p(ptr), _p(col), _p(val), _p(lp), _p(lc), _p(lv))
    if st != 0:
        raise ArithmeticError(f"IC(0) breakdown at row {-st - 1}")
    return lp, lc[:nl], lv[:nl]


def ic0_solve(lp, lc, lv, b):
    n = len(lp) - 1
    x = np.zeros(n)
    lib().orc_ic0_solve(n, _p(_i(lp)), _p(_i(lc)), _p(_d(lv)), _p(_d(b)), _p(x))
    return x


def pcg(ptr, col, val, b, tol_abs, max_iter=100000, factor=None, x0=None):
    """Jacobi (factor None) or IC(0) PCG, FP64.  Returns (x, iters)."""
    n = len(ptr) - 1
    x = np.zeros(n)
    if factor is None:
        lp = lc = lv = None
        pc = 0
    else:
        lp, lc, lv = (_i(factor[0]), _i(factor[1]), _d(factor[2]))
        pc = 1
    it = lib().orc_pcg(n, _p(_i(ptr)), _p(_i(col)), _p(_d(val)), pc,
                       _p(lp) if pc else None, _p(lc) if pc else None, _p(lv) if pc else None,
                       _p(_d(b)), _p(_d(x0)) if x0 is not None else None, max_iter, C.c_double(tol_abs),
                       _p(x))
    return x, it


def pcg_many(ptr, col, val, B, tol_rel, factor=None, threads=1, max_iter=100000):
    """Every column of B (n x k) solved by `pcg` to tol_rel * ||b_c||, the
    columns spread over `threads` threads.  Returns (X, total iterations)."""
    n = len(ptr) - 1
    B = np.ascontiguousarray(B, dtype=np.float64).reshape(n, -1)
    X = np.zeros_like(B)
    if factor is None:
        lp = lc = lv = None
        pc = 0
    else:
        lp, lc, lv = (_i(factor[0]), _i(factor[1]), _d(factor[2]))
        pc = 1
    it = lib().orc_pcg_many(n, _p(_i(ptr)), _p(_i(col)), _p(_d(val)), pc,
                            _p(lp) if pc else None, _p(lc) if pc else None,
                            _p(lv) if pc else None, B.shape[1], _p(B), max_iter,
                            C.c_double(tol_rel), _p(X), threads)
    return X, int(it)


def permute_csr(ptr, col, val, perm):
    n = len(ptr) - 1
    optr = np.zeros(n + 1, np.int32)
    ocol = np.zeros(len(col), np.int32)
    oval = np.zeros(len(col))
    lib().orc_permute_csr(n, _p(_i(ptr)), _p(_i(col)), _p(_d(val)), _p(_i(perm)), _p(optr),
                          _p(ocol), _p(oval))
    return optr, ocol, oval


def greedy_color(ptr, col):
    n = len(ptr) - 1
    color = np.zeros(n, np.int32)
    nc = lib().orc_greedy_color(n, _p(_i(ptr)), _p(_i(col)), _p(color))
    return color, nc


def color_classes(color, ncolors):
    """Rows of each colour, ascending: (color_ptr[ncolors + 1], color_rows)."""
    color = np.asarray(color)
    rows = np.argsort(color, kind="stable").astype(np.int32)
    cptr = np.zeros(ncolors + 1, np.int32)
    cptr[1:] = np.cumsum(np.bincount(color, minlength=ncolors))
    return cptr, rows


def csr_diag(ptr, col, val):
    """A_ii from CSR (0 where the row stores no diagonal)."""
    n = len(ptr) - 1
    d = np.zeros(n)
    for i in range(n):
        for q in range(ptr[i], ptr[i + 1]):
            if col[q] == i:
                d[i] = val[q]
    return d


def sgs_sweep(ptr, col, val, b, x0, sweeps=1, dtype=np.float64, color=None):
    """SgsColor (SgsColor.lean:87-130): `sweeps` symmetric colour sweeps of
    x (n or n x k) for A x = b, greedy colours unless `color` = (color,
    ncolors) is given.  FP64 or FP32 arithmetic (dtype)."""
    n = len(ptr) - 1
    if color is None:
        color = greedy_color(ptr, col)
    cptr, rows = color_classes(*color)
    b2 = np.asarray(b).reshape(n, -1)
    k = b2.shape[1]
    if dtype == np.float32:
        v = np.ascontiguousarray(val, np.float32)
        d = np.ascontiguousarray(csr_diag(ptr, col, val), np.float32)
        bb = np.ascontiguousarray(b2, np.float32)
        x = np.ascontiguousarray(np.asarray(x0).reshape(n, k), np.float32).copy()
        fn = lib().orc_sgs_sweep_f32
    else:
        v, d, bb = _d(val), _d(csr_diag(ptr, col, val)), _d(b2)
        x = _d(np.asarray(x0).reshape(n, k)).copy()
        fn = lib().orc_sgs_sweep_f64
    fn(n, _p(_i(ptr)), _p(_i(col)), _p(v), _p(d), color[1], _p(cptr), _p(rows), k, _p(bb),
       _p(x), sweeps)
    return x.reshape(np.asarray(b).shape)


def sgs_apply(ptr, col, val, r, color=None):
    """SGS preconditioner: one symmetric sweep from x = 0 for b = r."""
    return sgs_sweep(ptr, col, val, r, np.zeros_like(np.asarray(r, np.float64)), 1,
                     color=color)


def restrict_sum(parent, n_coarse, fine, dtype=np.float64):
    """HierarchicalSparsify restriction (HierarchicalSparsify.lean:64-93):
    coarse[c] = sum of fine[i] over parent[i] == c, i ascending."""
    f = np.asarray(fine)
    f2 = f.reshape(f.shape[0], -1)
    k = f2.shape[1]
    if dtype == np.float32:
        ff = np.ascontiguousarray(f2, np.float32)
        out = np.zeros((n_coarse, k), np.float32)
        fn = lib().orc_restrict_sum_f32
    else:
        ff = _d(f2)
        out = np.zeros((n_coarse, k))
        fn = lib().orc_restrict_sum_f64
    fn(len(parent), _p(_i(parent)), n_coarse, k, _p(ff), _p(out))
    return out.reshape((n_coarse,) + f.shape[1:])


# -------------------------------------------------------------------- dense --
def dense_solve(A, b):
    A = _d(A)
    n = A.shape[0]
    x = np.zeros(n)
    lib().orc_dense_solve(n, _p(A), _p(_d(b)), _p(x))
    return x


def dense_solve_multi(A, B):
    A, B = _d(A), _d(B)
    n, k = B.shape
    X = np.zeros((n, k))
    lib().orc_dense_solve_multi(n, k, _p(A), _p(B), _p(X))
    return X


def lu_factor(A):
    A = _d(A)
    n = A.shape[0]
    a = np.zeros_like(A)
    piv = np.zeros(n, np.int32)
    lib().orc_lu_factor(n, _p(A), _p(a), _p(piv))
    return a, piv


def lu_solve(a, piv, b):
    n = a.shape[0]
    x = np.zeros(n)
    lib().orc_lu_solve(n, _p(_d(a)), _p(_i(piv)), _p(_d(b)), _p(x))
    return x


def cholesky_factor(A):
    A = _d(A)
    L = np.zeros_like(A)
    lib().orc_cholesky_factor(A.shape[0], _p(A), _p(L))
    return L


def cholesky_solve(L, b):
    L = _d(L)
    x = np.zeros(L.shape[0])
    lib().orc_cholesky_solve(L.shape[0], _p(L), _p(_d(b)), _p(x))
    return x


# ---------------------------------------------------------------------- HSC --
def hsc_eliminate_vertex(edges, v):
    """edges: list of (u, w, weight) with u < w."""
    e = list(edges)
    ne = len(e)
    eu = _i([a for a, _, _ in e] or [0])
    ev = _i([b for _, b, _ in e] or [0])
    ew = _d([c for _, _, c in e] or [0.0])
    cap = ne + ne * ne + 4
    ou, ov, ow = np.zeros(cap, np.int32), np.zeros(cap, np.int32), np.zeros(cap)
    m = lib().orc_hsc_eliminate_vertex(ne, _p(eu), _p(ev), _p(ew), v, _p(ou), _p(ov), _p(ow))
    return [(int(ou[i]), int(ov[i]), float(ow[i])) for i in range(m)]


def hsc_edge_weight(edges, a, b):
    e = list(edges)
    eu = _i([x for x, _, _ in e] or [0])
    ev = _i([y for _, y, _ in e] or [0])
    ew = _d([z for _, _, z in e] or [0.0])
    return lib().orc_hsc_edge_weight(len(e), _p(eu), _p(ev), _p(ew), a, b)


def hsc_degree_weight(edges, v):
    e = list(edges)
    eu = _i([x for x, _, _ in e] or [0])
    ev = _i([y for _, y, _ in e] or [0])
    ew = _d([z for _, _, z in e] or [0.0])
    return lib().orc_hsc_degree_weight(len(e), _p(eu), _p(ev), _p(ew), v)


# ---------------------------------------------------------------- adjoint --
def adjoint(ff, fc, g, x=None, xc=None, tol_rel=1e-14, factor=None):
    """Adjoint of the harmonic solve A_ff x_f = -A_fc x_c (builder-defined;
    the reference only runs HarmonicSolve.solveMulti forward,
    HarmonicSolve.lean:51-66).  ff = (ptr, col, val) of A_ff; fc = (ptr,
    col, val, n_con) of A_fc or None; g = dL/dx_f (n x k).
      lam = A_ff^-1 g (FP64 PCG per column, IC(0) when `factor` is given);
      gc  = -A_fc^T lam (n_con x k);
      gw  = dL/dw per stored entry of A_ff then A_fc (needs x, xc), w the
            weight of the edge the entry couples: -sum_cols (lam_i - lam_j)
            (x_i - x_j), lam = 0 on constrained vertices, 0 on the diagonal.
    Returns (lam, gc, gw)."""
    ptr, col, val = (np.asarray(a) for a in ff)
    g = np.asarray(g, dtype=np.float64).reshape(len(ptr) - 1, -1)
    n, k = g.shape
    lam = np.zeros_like(g)
    for c in range(k):
        nb = np.linalg.norm(g[:, c])
        if nb > 0:
            lam[:, c], _ = pcg(ptr, col, val, g[:, c], tol_rel * nb, factor=factor)
    rows = np.repeat(np.arange(n), np.diff(ptr))
    gc = None
    if fc is not None:
        fptr, fcol, fval, ncon = fc
        frows = np.repeat(np.arange(n), np.diff(fptr))
        gc = np.zeros((ncon, k))
        np.add.at(gc, np.asarray(fcol), -np.asarray(fval)[:, None] * lam[frows])
    gw = None
    if x is not None:
        x = np.asarray(x, dtype=np.float64).reshape(n, k)
        d = -((lam[rows] - lam[col]) * (x[rows] - x[col])).sum(1)
        d[rows == col] = 0.0
        parts = [d]
        if fc is not None:
            xc = np.asarray(xc, dtype=np.float64).reshape(ncon, k)
            parts.append(-(lam[frows] * (x[frows] - xc[np.asarray(fcol)])).sum(1))
        gw = np.concatenate(parts)
    return lam, gc, gw


# -------------------------------------------------------------------- HSC --
def hsc_build(ptr, col, val, theta=0.015, coarse_max=512, max_levels=24, min_frac=0.05,
              comp="lump"):
    """Hierarchical Schur elimination with sparsification (after
    Krishnan-Fattal-Szeliski 2013; the per-vertex rule is
    HierarchicalSparsifyCompensate.eliminateVertex,
    HierarchicalSparsifyCompensate.lean:96-125, generalised from a graph
    Laplacian to an SPD matrix: the pivot is the diagonal, not the degree).

    Level l, matrix A (CSR, sorted columns):
      1. F = greedy independent set, vertices visited by (off-diagonal count,
         index) ascending; C = the rest.  Stop (A is the coarse matrix) when
         n <= coarse_max, at max_levels, or when |F| < min_frac n.
      2. S = A_CC - A_CF D_F^-1 A_FC exactly: starting from A_CC, for every
         f in F ascending, for every pair (c, c') of f's neighbours
         (ascending, c' over all of them), S[c][c'] -= a_cf a_fc' / a_ff.
      3. Off-diagonal entries of S outside the pattern of A_CC (fill) with
         |s_ij| < theta sqrt(s_ii s_jj) (tested on S before any change) are
         dropped; each adds |s_ij| to s_ii (rows ascending, columns
         ascending): S + |s|(e_i e_i^T + e_j e_j^T) - s(e_i e_j^T + e_j e_i^T)
         >= S, so every level stays SPD.
    `comp` = "lump" adds s_ij itself instead (s (e_i - e_j)(e_i - e_j)^T:
    row sums, hence the action on constants, are kept; far fewer iterations
    on Laplacians, definite while theta is small).  Drops are applied in
    (i, j) ascending order.
    Returns a dict of levels (F, C, dinvF, R, P as CSR) and the dense coarse
    matrix.  Indices are level-local; C keeps ascending order."""
    ptr = np.asarray(ptr, np.int64)
    col = np.asarray(col, np.int64)
    val = np.asarray(val, np.float64)
    rows = [dict(zip(col[ptr[i]:ptr[i + 1]].tolist(), val[ptr[i]:ptr[i + 1]].tolist()))
            for i in range(len(ptr) - 1)]
    levels = []
    while True:
        n = len(rows)
        if n <= coarse_max or len(levels) >= max_levels:
            break
        deg = [len(r) - (1 if i in r else 0) for i, r in enumerate(rows)]
        order = sorted(range(n), key=lambda i: (deg[i], i))
        isF = [False] * n
        blocked = [False] * n
        for v in order:
            if not blocked[v]:
                isF[v] = True
                blocked[v] = True
                for u in rows[v]:
                    blocked[u] = True
        F = [i for i in range(n) if isF[i]]
        if len(F) < min_frac * n:
            break
        C = [i for i in range(n) if not isF[i]]
        cidx = {c: t for t, c in enumerate(C)}
        fidx = {f: t for t, f in enumerate(F)}
        S = [{cidx[j]: a for j, a in rows[c].items() if not isF[j]} for c in C]
        orig = [set(r.keys()) for r in S]
        for f in F:
            d = rows[f][f]
            nb = sorted(j for j in rows[f] if j != f)
            for c in nb:
                acf = rows[c][f]
                rc = S[cidx[c]]
                for c2 in nb:
                    t = cidx[c2]
                    rc[t] = rc.get(t, 0.0) - acf * rows[f][c2] / d
        drop = []
        for i, r in enumerate(S):
            for j in sorted(r):
                s = r[j]
                if j > i and j not in orig[i] and abs(s) < theta * math.sqrt(r[i] * S[j][j]):
                    drop.append((i, j))
        for i, j in drop:
            s = S[i][j]
            del S[i][j]
            del S[j][i]
            w = abs(s) if comp == "diag" else s
            S[i][i] += w
            S[j][j] += w
        dinv = np.array([1.0 / rows[f][f] for f in F])
        Rp, Rc, Rv = [0], [], []
        for c in C:
            ent = sorted((fidx[j], -a / rows[j][j]) for j, a in rows[c].items() if isF[j])
            Rc += [e[0] for e in ent]
            Rv += [e[1] for e in ent]
            Rp.append(len(Rc))
        Pp, Pc, Pv = [0], [], []
        for f in F:
            ent = sorted((cidx[j], -a / rows[f][f]) for j, a in rows[f].items() if j != f)
            Pc += [e[0] for e in ent]
            Pv += [e[1] for e in ent]
            Pp.append(len(Pc))
        levels.append({"n": n, "F": np.array(F, np.int32), "C": np.array(C, np.int32),
                       "dinv": dinv,
                       "R": (np.array(Rp, np.int32), np.array(Rc, np.int32), np.array(Rv)),
                       "P": (np.array(Pp, np.int32), np.array(Pc, np.int32), np.array(Pv))})
        rows = [dict(sorted(r.items())) for r in S]
    n = len(rows)
    dense = np.zeros((n, n))
    for i, r in enumerate(rows):
        for j, a in r.items():
            dense[i, j] = a
    return {"levels": levels, "coarse": dense}


def hsc_apply(h, r):
    """z = M^-1 r for the hierarchy of hsc_build (FP64): per level
    r_C += R r_F (terms added to r_C in ascending F index) down to the coarse
    level, z_c = A_c^-1 r_c there (dense Cholesky), then
    z_F = dinv r_F + P z_C (terms added in ascending C index), z_C = z_c."""
    r = np.asarray(r, np.float64)
    vec = r.ndim == 1
    r = r.reshape(len(r), -1)
    rs = [r]
    for L in h["levels"]:
        cur = rs[-1]
        Rp, Rc, Rv = L["R"]
        nxt = cur[L["C"]].copy()
        for c in range(len(L["C"])):
            for e in range(Rp[c], Rp[c + 1]):
                nxt[c] = nxt[c] + Rv[e] * cur[L["F"][Rc[e]]]
        rs.append(nxt)
    Lc = np.linalg.cholesky(h["coarse"]) if len(h["coarse"]) else None
    z = rs[-1].copy()
    if Lc is not None:
        import scipy.linalg as sla
        z = sla.cho_solve((Lc, True), rs[-1])
    for li in range(len(h["levels"]) - 1, -1, -1):
        L = h["levels"][li]
        cur = rs[li]
        Pp, Pc, Pv = L["P"]
        zl = np.zeros_like(cur)
        zl[L["C"]] = z
        for f in range(len(L["F"])):
            acc = L["dinv"][f] * cur[L["F"][f]]
            for e in range(Pp[f], Pp[f + 1]):
                acc = acc + Pv[e] * z[Pc[e]]
            zl[L["F"][f]] = acc
        z = zl
    return z.reshape(-1) if vec else z
