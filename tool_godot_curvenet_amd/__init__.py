"""MI355X (gfx950) profile-curves linear-algebra core.

Drop-in for the solver path of V-Sekai-fire/TOOL_godot_curvenet
(lean/Curvenet: HalfedgeBuilder, CutMesh, CutMeshLaplacian, SparseLinAlg,
IncompleteCholesky, HarmonicSolve): hand-written HIP kernels behind the C
ABI of ``include/curvenet_la.h``; this package is the Python host side.
"""
from . import _lib, autotune, curves, ddm, formats
from .curvenet import (  # noqa: F401
    HSC,
    IC0,
    SGS,
    Restriction,
    CutLaplacian,
    CsrMatrix,
    CsrSystem,
    BatchSystem,
    cg,
    cg_with_guess,
    bezier3,
    cut_seams,
    deform_solve,
    deformation_gradients,
    dense_cholesky,
    dense_cholesky_solve,
    from_triangles,
    harmonic_solve,
    project_to_vertices,
    promote_vertex_samples,
    segment_gradients,
    spmm,
    spmv,
)
from .autograd import harmonic_solve as harmonic_solve_ad  # noqa: F401

__all__ = [
    "HSC", "SGS", "Restriction", "dense_cholesky", "dense_cholesky_solve", "IC0", "CutLaplacian", "CsrMatrix", "CsrSystem", "BatchSystem", "cg", "cg_with_guess", "cut_seams", "deform_solve", "bezier3", "segment_gradients", "deformation_gradients",
    "from_triangles", "harmonic_solve", "harmonic_solve_ad", "project_to_vertices", "promote_vertex_samples", "spmm",
    "spmv", "curves", "ddm", "formats", "autotune",
]
