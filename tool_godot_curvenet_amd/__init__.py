"""MI355X (gfx950) profile-curves linear-algebra core.

Drop-in for the solver path of V-Sekai-fire/TOOL_godot_curvenet
(lean/Curvenet: HalfedgeBuilder, CutMesh, CutMeshLaplacian, SparseLinAlg,
IncompleteCholesky, HarmonicSolve): hand-written HIP kernels behind the C
ABI of ``include/curvenet_la.h``; this package is the Python host side.
"""
from . import _lib
from .curvenet import (  # noqa: F401
    IC0,
    CutLaplacian,
    CsrMatrix,
    cg,
    cg_with_guess,
    cut_seams,
    from_triangles,
    harmonic_solve,
    project_to_vertices,
    spmm,
    spmv,
)

__all__ = [
    "IC0", "CutLaplacian", "CsrMatrix", "cg", "cg_with_guess", "cut_seams",
    "from_triangles", "harmonic_solve", "project_to_vertices", "spmm", "spmv",
]
