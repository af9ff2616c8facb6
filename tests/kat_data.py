"""Small instances from the reference's own native_decide examples,
transcribed as data (file:line of each source in the comments)."""
import math

import numpy as np

S3 = 0.8660254037844386

# Halfedge.lean:137-200 -- hand-built meshes (target, twin, next, face), -1 = none
TRIANGLE = dict(nv=3, nf=1, target=[1, 2, 0, 2, 1, 0], twin=[5, 4, 3, 2, 1, 0],
                next=[1, 2, 0, 4, 5, 3], face=[0, 0, 0, -1, -1, -1])
QUAD = dict(nv=4, nf=1, target=[1, 2, 3, 0, 3, 2, 1, 0], twin=[7, 6, 5, 4, 3, 2, 1, 0],
            next=[1, 2, 3, 0, 5, 6, 7, 4], face=[0, 0, 0, 0, -1, -1, -1, -1])
TWO_TRI_STRIP = dict(nv=4, nf=2, target=[1, 2, 0, 1, 3, 2, 2, 0, 3, 1],
                     twin=[7, 3, 6, 1, 9, 8, 2, 0, 5, 4],
                     next=[1, 2, 0, 4, 5, 3, 8, 6, 9, 7],
                     face=[0, 0, 0, 1, 1, 1, -1, -1, -1, -1])


def he(d):
    out = {k: np.asarray(d[k], np.int32) for k in ("target", "twin", "next", "face")}
    out["nv"], out["nf"], out["nh"] = d["nv"], d["nf"], len(d["target"])
    return out


# HalfedgeBuilder.lean:106-112
SINGLE_TRI = (3, [(0, 1, 2)])
STRIP = (4, [(0, 1, 2), (2, 1, 3)])
TET = (4, [(0, 1, 2), (0, 2, 3), (0, 3, 1), (1, 3, 2)])

# PolygonLaplacian.lean:111-131 / CutMeshLaplacian.lean:108-156
EQ_TRI = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.5, 0.0, S3]])
UNIT_SQUARE = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [1.0, 0.0, 1.0], [0.0, 0.0, 1.0]])
PENTAGON = np.array([[math.cos(6.283185307179586 * k / 5.0), 0.0,
                      math.sin(6.283185307179586 * k / 5.0)] for k in range(5)])
STRIP_POS = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.5, 0.0, S3], [1.5, 0.0, S3]])

# SparseLinAlg.lean:176-240 (CSR: ptr, col, val)
IDENTITY3 = ([0, 1, 2, 3], [0, 1, 2], [1.0, 1.0, 1.0])
DIAG234 = ([0, 1, 2, 3], [0, 1, 2], [2.0, 3.0, 4.0])
SPD2 = ([0, 2, 4], [0, 1, 0, 1], [2.0, 1.0, 1.0, 2.0])
LAPLACIAN4 = ([0, 2, 5, 8, 10], [0, 1, 0, 1, 2, 1, 2, 3, 2, 3],
              [2.0, -1.0, -1.0, 2.0, -1.0, -1.0, 2.0, -1.0, -1.0, 2.0])
LAPLACIAN4_DENSE = np.array([[2.0, -1, 0, 0], [-1, 2, -1, 0], [0, -1, 2, -1], [0, 0, -1, 2]])
# Features/KnownKnowns/SolverKernels.lean:9-18
TRI3 = ([0, 2, 5, 7], [0, 1, 0, 1, 2, 1, 2], [4.0, -1.0, -1.0, 4.0, -1.0, -1.0, 4.0])

# IncompleteCholesky.lean:136-180
ATRI = np.array([[4.0, -1.0, 0.0], [-1.0, 4.0, -1.0], [0.0, -1.0, 4.0]])
FULLPAT = np.array([[0.0, 1, 1], [1, 0, 1], [1, 1, 0]])
AWIDE = np.array([[1.0e-2, 1.0e-3, 0.0], [1.0e-3, 1.0, 1.0e-2], [0.0, 1.0e-2, 1.0e+4]])
WIDEPAT = np.array([[0.0, 1, 0], [1, 0, 1], [0, 1, 0]])

# HierarchicalSparsifyCompensate.lean:143-200
PATH3 = [(0, 1, 1.0), (1, 2, 2.0)]
STAR4 = [(0, 1, 1.0), (0, 2, 2.0), (0, 3, 3.0)]
TRI_PLUS_LEAF = [(0, 1, 1.0), (0, 2, 1.0), (1, 2, 1.0), (0, 3, 4.0)]

# GraphColoring.lean:84-113 (adjacency lists)
COLOR_GRAPHS = {
    "triangle": ([[1, 2], [0, 2], [0, 1]], 3),
    "cycle4": ([[1, 3], [0, 2], [1, 3], [0, 2]], 2),
    "path5": ([[1], [0, 2], [1, 3], [2, 4], [3]], 2),
    "star4": ([[1, 2, 3, 4], [0], [0], [0], [0]], 2),
}


def adj_to_csr(adj):
    ptr = [0]
    col = []
    for row in adj:
        col.extend(sorted(row))
        ptr.append(len(col))
    return np.asarray(ptr, np.int32), np.asarray(col, np.int32)


def csr_np(m):
    return (np.asarray(m[0], np.int32), np.asarray(m[1], np.int32), np.asarray(m[2], np.float64))


def dense_to_csr(A, keep=None):
    A = np.asarray(A, dtype=np.float64)
    n = A.shape[0]
    ptr, col, val = [0], [], []
    for i in range(n):
        for j in range(n):
            if A[i, j] != 0.0 or i == j or (keep is not None and keep[i, j]):
                col.append(j)
                val.append(A[i, j])
        ptr.append(len(col))
    return np.asarray(ptr, np.int32), np.asarray(col, np.int32), np.asarray(val)
