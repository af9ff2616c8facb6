"""Python host side of the gfx950 library, named after the spec functions.

Every call goes through the C ABI (``_lib``); device memory is held in torch
CUDA tensors (plumbing only).  Functions mirror the reference's Lean
definitions:

=========================  ==================================================
``from_triangles``         HalfedgeBuilder.fromTriangles (HalfedgeBuilder.lean:34-90)
``cut_seams``              seam duplication (cn_cut_seams; builder-defined rule)
``project_to_vertices``    SurfaceProjection.projectToVertices (:46-64)
``promote_vertex_samples`` SurfaceProjection.promoteVertexSamples (:64-80)
``CutLaplacian``           CutMeshLaplacian.assembleVtLhV (:86-93), split A_ff / A_fc
``CutLaplacian.harmonic_rhs``  HarmonicSolve.solveMulti RHS (:35-45, 57-64)
``spmv`` / ``spmm``        SparseLinAlg.spmv (:44-53) / SpmvMulti
``cg`` / ``cg_with_guess`` SparseLinAlg.cg / cgWithGuess (:104-163)
``IC0``                    IncompleteCholesky (CSR no-fill, :26-101)
``harmonic_solve``         HarmonicSolve.solveMulti (:51-66) on the sparse path
=========================  ==================================================
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as L


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("no HIP device visible: the gfx950 library needs a GPU "
                           "(there is no CPU fallback)")
    return torch


def _dev(a, dtype):
    torch = _torch()
    t = torch.as_tensor(np.ascontiguousarray(a))
    return t.to(device="cuda", dtype=dtype).contiguous()


def _stream():
    return C.c_void_p(_torch().cuda.current_stream().cuda_stream)


def _p(t):
    return C.c_void_p(t.data_ptr()) if t is not None and t.numel() else C.c_void_p(0)


def _np(ptr: int, count: int, dtype) -> np.ndarray:
    dt = np.dtype(dtype)
    _torch().cuda.synchronize()
    return np.frombuffer(L.d2h(ptr, count * dt.itemsize), dtype=dt).copy()


def _as_tensor(x, dtype):
    torch = _torch()
    if isinstance(x, torch.Tensor):
        return x.to(device="cuda", dtype=dtype).contiguous()
    return _dev(x, dtype)


# ------------------------------------------------------------------ topology --
def from_triangles(nv: int, tris) -> dict:
    """HalfedgeBuilder.fromTriangles on the GPU; returns numpy arrays
    target / twin / next / face (-1 = none)."""
    torch = _torch()
    lib = L.load()
    t = _dev(np.asarray(tris, dtype=np.int32).reshape(-1, 3), torch.int32)
    nf = t.shape[0]
    cap = max(6 * nf, 1)
    bufs = [torch.empty(cap, dtype=torch.int32, device="cuda") for _ in range(4)]
    nh = C.c_int32(0)
    L.check(lib.cn_halfedge_from_triangles(nv, nf, _p(t), *(_p(b) for b in bufs),
                                           C.byref(nh), _stream()), "from_triangles")
    out = {k: b[:nh.value].cpu().numpy() for k, b in zip(("target", "twin", "next", "face"), bufs)}
    out["nv"], out["nf"], out["nh"] = nv, nf, nh.value
    return out


def cut_seams(nv: int, tris, cut_edges):
    """Cut the mesh open along the given edges.  Returns (new_tris, origin,
    constrained, nv_after)."""
    torch = _torch()
    lib = L.load()
    t = _dev(np.asarray(tris, dtype=np.int32).reshape(-1, 3), torch.int32)
    e = _dev(np.asarray(cut_edges, dtype=np.int32).reshape(-1, 2), torch.int32)
    nf, ncut = t.shape[0], e.shape[0]
    cap = nv + 2 * ncut + 1
    nt = torch.empty_like(t)
    origin = torch.empty(cap, dtype=torch.int32, device="cuda")
    con = torch.empty(cap, dtype=torch.uint8, device="cuda")
    nv2 = C.c_int32(0)
    L.check(lib.cn_cut_seams(nv, nf, _p(t), ncut, _p(e), _p(nt), _p(origin), _p(con),
                             C.byref(nv2), _stream()), "cut_seams")
    n2 = nv2.value
    return (nt.cpu().numpy(), origin[:n2].cpu().numpy(), con[:n2].cpu().numpy(), n2)


def project_to_vertices(knots, pos) -> np.ndarray:
    torch = _torch()
    lib = L.load()
    q = _dev(np.asarray(knots, dtype=np.float64).reshape(-1, 3), torch.float64)
    p = _dev(np.asarray(pos, dtype=np.float64).reshape(-1, 3), torch.float64)
    out = torch.empty(q.shape[0], dtype=torch.int32, device="cuda")
    L.check(lib.cn_project_to_vertices(p.shape[0], _p(p), q.shape[0], _p(q), _p(out),
                                       _stream()), "project_to_vertices")
    return out.cpu().numpy()


def promote_vertex_samples(vertex_count: int, index):
    """SurfaceProjection.promoteVertexSamples (SurfaceProjection.lean:64-80)
    on the device: the sample column of every knot, in encounter order.
    Returns (knot_to_col, vertex_to_col, ncols); -1 marks knots without a
    vertex and vertices no knot landed on."""
    torch = _torch()
    lib = L.load()
    idx = _dev(np.asarray(index, dtype=np.int32).reshape(-1), torch.int32)
    kc = torch.empty(idx.shape[0], dtype=torch.int32, device="cuda")
    vc = torch.empty(int(vertex_count), dtype=torch.int32, device="cuda")
    nc = C.c_int32(0)
    L.check(lib.cn_promote_vertex_samples(int(vertex_count), idx.shape[0], _p(idx), _p(kc),
                                          _p(vc), C.byref(nc), _stream()),
            "promote_vertex_samples")
    return kc.cpu().numpy(), vc.cpu().numpy(), nc.value


# ----------------------------------------------------------------- matrices --
@dataclass
class CsrMatrix:
    """SparseLinAlg.SparseMatrixCSR (SparseLinAlg.lean:35-41)."""
    rows: int
    cols: int
    row_ptr: np.ndarray
    col_idx: np.ndarray
    values: np.ndarray

    @property
    def nnz(self) -> int:
        return int(self.row_ptr[-1])

    def dense(self) -> np.ndarray:
        d = np.zeros((self.rows, self.cols))
        for i in range(self.rows):
            a, b = self.row_ptr[i], self.row_ptr[i + 1]
            d[i, self.col_idx[a:b]] += self.values[a:b]
        return d

    def device(self):
        torch = _torch()
        return (_dev(self.row_ptr.astype(np.int32), torch.int32),
                _dev(self.col_idx.astype(np.int32), torch.int32),
                _dev(self.values.astype(np.float64), torch.float64))


def spmm(A: CsrMatrix, X, precision: str = "f64"):
    """Y = A X for an n x k (or n) block; f64 is the spec-exact form."""
    torch = _torch()
    lib = L.load()
    X = np.asarray(X) if not isinstance(X, torch.Tensor) else X
    vec = X.ndim == 1
    k = 1 if vec else X.shape[1]
    ptr, col, val = A.device()
    if precision == "f64":
        x = _as_tensor(X, torch.float64)
        y = torch.empty((A.rows, k), dtype=torch.float64, device="cuda")
        L.check(lib.cn_spmm_f64(A.rows, _p(ptr), _p(col), _p(val), k, _p(x), _p(y),
                                _stream()), "spmm_f64")
    else:
        x = _as_tensor(X, torch.float32)
        v32 = val.float()
        y = torch.empty((A.rows, k), dtype=torch.float32, device="cuda")
        L.check(lib.cn_spmm_f32(A.rows, _p(ptr), _p(col), _p(v32), k, _p(x), _p(y),
                                _stream()), "spmm_f32")
    y = y.cpu().numpy()
    return y.reshape(-1) if vec else y


def spmv(A: CsrMatrix, x, precision: str = "f64"):
    return spmm(A, np.asarray(x).reshape(-1), precision)


# ------------------------------------------------------------------ systems --
class _System:
    """Owner of a cn_system handle."""

    # mixed precision: a residual replacement once the iterated ||r|| has
    # fallen by this factor (cn_pcg_opts.inner_tol), unless the autotuner set
    # another for this system or the caller passes one
    INNER_TOL = 0.0316

    def __init__(self, handle: int):
        self._h = C.c_void_p(handle)
        self.inner_tol = self.INNER_TOL
        info = L.SystemInfo()
        L.check(L.load().cn_system_get_info(self._h, C.byref(info)), "system_get_info")
        self.info = info

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                L.load().cn_system_destroy(h)
            except Exception:
                pass
            self._h = C.c_void_p(0)

    @property
    def handle(self):
        return self._h

    @property
    def n_free(self) -> int:
        return self.info.n_free

    TUNE_FIELDS = ("spmm_form", "spmm_gens", "level_cap", "tail_rows", "chain_levels",
                   "coarse_prec", "down_cap")

    def set_tune(self, tune: dict | None):
        """Launch shapes of the solver kernels (cn_system_set_tune): a dict of
        spmm_form, spmm_gens, level_cap, tail_rows, chain_levels, coarse_prec, down_cap;
        missing keys (or None) keep the built-in choice."""
        t = L.Tune()
        for f in self.TUNE_FIELDS:
            setattr(t, f, int((tune or {}).get(f, 0)))
        L.check(L.load().cn_system_set_tune(self._h, C.byref(t)), "system_set_tune")

    def get_tune(self) -> dict:
        t = L.Tune()
        L.check(L.load().cn_system_get_tune(self._h, C.byref(t)), "system_get_tune")
        return {f: getattr(t, f) for f in self.TUNE_FIELDS}

    def A_ff(self) -> CsrMatrix:
        i = self.info
        n = i.n_free
        return CsrMatrix(n, n, _np(i.d_ff_ptr, n + 1, np.int32), _np(i.d_ff_col, i.nnz_ff, np.int32),
                         _np(i.d_ff_val, i.nnz_ff, np.float64))

    @staticmethod
    def _opts(precond, precision, tol, max_iter, inner_tol, check_every, max_refine, warm,
              graph=True):
        o = L.PcgOpts()
        L.load().cn_pcg_default_opts(C.byref(o))
        o.graph = int(bool(graph))
        o.precond = {"ic0": L.CN_PRECOND_IC0, "jacobi": L.CN_PRECOND_JACOBI}[precond]
        o.precision = {"mixed": L.CN_PREC_MIXED, "fp64": L.CN_PREC_FP64,
                       "fp64_plain": L.CN_PREC_FP64_PLAIN}[precision]
        o.tol_rel, o.max_iter, o.inner_tol = tol, max_iter, inner_tol
        o.check_every, o.max_refine, o.warm_start = check_every, max_refine, int(warm)
        return o

    def solve(self, b, precond: str = "ic0", precision: str = "mixed", tol: float = 1e-9,
              x0=None, ic0: "IC0 | None" = None, max_iter: int = 20000,
              inner_tol: float | None = None, check_every: int = 8, max_refine: int = 30,
              hsc: "HSC | None" = None, sgs: "SGS | None" = None, graph: bool = True):
        """Solve A_ff X = B (n x k or n); precond "ic0", "jacobi", "hsc" or
        "sgs".  graph: replay the inner iterations as a captured hipGraph
        (the same results as direct launches).  inner_tol: the mixed
        precision residual-replacement factor (None: this system's, see
        INNER_TOL and autotune).  Returns (x, SolveInfo)."""
        if inner_tol is None:
            inner_tol = self.inner_tol
        torch = _torch()
        lib = L.load()
        bt = _as_tensor(b, torch.float64)
        vec = bt.dim() == 1
        bt = bt.reshape(self.n_free, -1).contiguous()
        k = bt.shape[1]
        x = (_as_tensor(x0, torch.float64).reshape(self.n_free, k).clone()
             if x0 is not None else torch.zeros_like(bt))
        o = self._opts("jacobi" if precond in ("hsc", "sgs") else precond, precision, tol,
                       max_iter, inner_tol, check_every, max_refine, x0 is not None, graph)
        info = L.SolveInfo()
        if precond == "hsc":
            if hsc is None:
                hsc = HSC(self)
            st = lib.cn_pcg_solve_hsc(self._h, hsc.handle, C.byref(o), k, _p(bt), _p(x),
                                      C.byref(info), _stream())
        elif precond == "sgs":
            if sgs is None:
                sgs = SGS(self)
            st = lib.cn_pcg_solve_sgs(self._h, sgs.handle, C.byref(o), k, _p(bt), _p(x),
                                      C.byref(info), _stream())
        else:
            if precond == "ic0" and ic0 is None:
                ic0 = IC0(self)
            st = lib.cn_pcg_solve(self._h, ic0.handle if ic0 is not None else C.c_void_p(0),
                                  C.byref(o), k, _p(bt), _p(x), C.byref(info), _stream())
        L.check(st, "pcg_solve", ok=(L.CN_OK, L.CN_MAXIT, L.CN_BREAKDOWN))
        return (x.reshape(-1) if vec else x), info

    def adjoint(self, g, x=None, xc=None, precond: str = "ic0", precision: str = "mixed",
                tol: float = 1e-9, ic0: "IC0 | None" = None, weights: bool = False,
                max_iter: int = 20000, inner_tol: float | None = None, check_every: int = 8,
                max_refine: int = 30, hsc: "HSC | None" = None):
        """Adjoint of the harmonic solve (cn_adjoint_solve, or
        cn_adjoint_solve_hsc with precond "hsc") for g = dL/dx_f
        (n_free x k).  Returns (lambda, g_c, g_w, SolveInfo) as device
        tensors: lambda = A_ff^-1 g, g_c = -A_fc^T lambda (None without
        constrained columns), g_w = dL/dw per stored entry of A_ff then A_fc
        (only with weights=True, which needs the forward x and x_c)."""
        if inner_tol is None:
            inner_tol = self.inner_tol
        torch = _torch()
        lib = L.load()
        gt = _as_tensor(g, torch.float64)
        vec = gt.dim() == 1
        gt = gt.reshape(self.n_free, -1).contiguous()
        k = gt.shape[1]
        i = self.info
        lam = torch.empty_like(gt)
        gc = (torch.empty((i.n_con, k), dtype=torch.float64, device="cuda")
              if i.n_con > 0 else None)
        gw = xt = xct = None
        if weights:
            xt = _as_tensor(x, torch.float64).reshape(self.n_free, k).contiguous()
            if i.n_con > 0:
                xct = _as_tensor(xc, torch.float64).reshape(i.n_con, k).contiguous()
            gw = torch.empty(i.nnz_ff + i.nnz_fc, dtype=torch.float64, device="cuda")
        o = self._opts("jacobi" if precond == "hsc" else precond, precision, tol, max_iter,
                       inner_tol, check_every, max_refine, False)
        info = L.SolveInfo()
        if precond == "hsc":
            if hsc is None:
                hsc = HSC(self)
            st = lib.cn_adjoint_solve_hsc(self._h, hsc.handle, C.byref(o), k, _p(gt), _p(xt),
                                          _p(xct), _p(lam), _p(gc), _p(gw), C.byref(info),
                                          _stream())
        else:
            if precond == "ic0" and ic0 is None:
                ic0 = IC0(self)
            st = lib.cn_adjoint_solve(self._h, ic0.handle if ic0 is not None else C.c_void_p(0),
                                      C.byref(o), k, _p(gt), _p(xt), _p(xct), _p(lam), _p(gc),
                                      _p(gw), C.byref(info), _stream())
        L.check(st, "adjoint_solve", ok=(L.CN_OK, L.CN_MAXIT, L.CN_BREAKDOWN))
        if vec:
            lam = lam.reshape(-1)
            gc = gc.reshape(-1) if gc is not None else None
        return lam, gc, gw, info


class CutLaplacian(_System):
    """VtLhV of a cut triangle mesh (CutMeshLaplacian.assembleVtLhV), stored
    as the free block A_ff and the coupling block A_fc."""

    def __init__(self, pos, tris, constrained, laplacian: str = "cot", delta: float = 0.0):
        """laplacian "cot" (the embedded cot, PolygonLaplacian) or "robust"
        (RobustLaplacian, mollified edge lengths with margin delta)."""
        torch = _torch()
        lib = L.load()
        self.pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
        self.tris = np.asarray(tris, dtype=np.int32).reshape(-1, 3)
        self.constrained = np.asarray(constrained, dtype=np.uint8).reshape(-1)
        t = _dev(self.tris, torch.int32)
        c = _dev(self.constrained, torch.uint8)
        p = _dev(self.pos, torch.float64)
        h = C.c_void_p(0)
        o = L.SystemOpts()
        lib.cn_system_default_opts(C.byref(o))
        o.laplacian = {"cot": L.CN_LAP_COT, "robust": L.CN_LAP_ROBUST}[laplacian]
        o.delta = delta
        L.check(lib.cn_system_create_ex(self.pos.shape[0], self.tris.shape[0], _p(t), _p(c),
                                        _p(p), C.byref(o), C.byref(h), _stream()),
                "system_create")
        super().__init__(h.value)

    @classmethod
    def from_mesh(cls, mesh, laplacian: str = "cot", delta: float = 0.0) -> "CutLaplacian":
        """Cut a generated mesh along its cut edges and build the system."""
        nt, origin, con, nv2 = cut_seams(mesh.nv, mesh.tris, mesh.cut_edges)
        pos = mesh.pos[origin]
        obj = cls(pos, nt, con, laplacian=laplacian, delta=delta)
        obj.origin = origin
        return obj

    @property
    def n_con(self) -> int:
        return self.info.n_con

    def A_fc(self) -> CsrMatrix:
        i = self.info
        n = i.n_free
        return CsrMatrix(n, i.n_con, _np(i.d_fc_ptr, n + 1, np.int32),
                         _np(i.d_fc_col, i.nnz_fc, np.int32), _np(i.d_fc_val, i.nnz_fc, np.float64))

    def maps(self):
        i = self.info
        return (_np(i.d_free_of, i.nv, np.int32), _np(i.d_con_of, i.nv, np.int32))

    def halfedges(self) -> dict:
        i = self.info
        return {k: _np(getattr(i, "d_he_" + k), i.nh, np.int32)
                for k in ("target", "twin", "next", "face")}

    def harmonic_rhs(self, Xc, as_numpy: bool = True):
        """-(Vt Lh C Xc) on the free rows, spec summation order (FP64)."""
        torch = _torch()
        lib = L.load()
        x = _as_tensor(Xc, torch.float64).reshape(self.n_con, -1).contiguous()
        k = x.shape[1]
        out = torch.empty((self.n_free, k), dtype=torch.float64, device="cuda")
        L.check(lib.cn_harmonic_rhs(self._h, k, _p(x), _p(out), _stream()), "harmonic_rhs")
        return out.cpu().numpy() if as_numpy else out

    def set_positions(self, pos):
        torch = _torch()
        self.pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
        p = _dev(self.pos, torch.float64)
        L.check(L.load().cn_system_set_positions(self._h, _p(p), _stream()), "set_positions")


class CsrSystem(_System):
    """A caller-assembled SPD matrix as a solver system."""

    def __init__(self, A: CsrMatrix):
        ptr, col, val = A.device()
        h = C.c_void_p(0)
        L.check(L.load().cn_system_from_csr(A.rows, _p(ptr), _p(col), _p(val), C.byref(h),
                                            _stream()), "system_from_csr")
        super().__init__(h.value)


def deform_solve(lap: "CutLaplacian", Fc, Xc, precond: str = "hsc", precision: str = "mixed",
                 tol: float = 1e-9, hsc: "HSC | None" = None, ic0: "IC0 | None" = None):
    """DeformSolve.solveDeformation (DeformSolve.lean:110-135) for P poses:
    Fc (n_con x 9P, a row-major 3x3 per pose), Xc (n_con x 3P).  Stage 1
    solves the free vertices' deformation gradients, stage 2 their
    positions.  Returns (x_free n_free x 3P, x_all nv x 3P with the
    constrained vertices overlaid (DeformSolve.lean:108-109), info1, info2)
    as device tensors."""
    torch = _torch()
    lib = L.load()
    fc = _as_tensor(Fc, torch.float64).reshape(lap.n_con, -1).contiguous()
    xc = _as_tensor(Xc, torch.float64).reshape(lap.n_con, -1).contiguous()
    P = fc.shape[1] // 9
    if fc.shape[1] != 9 * P or xc.shape[1] != 3 * P or P == 0:
        raise ValueError("Fc must be n_con x 9P and Xc n_con x 3P")
    kw = {"hsc": hsc} if precond == "hsc" else ({"ic0": ic0} if precond == "ic0" else {})
    b1 = lap.harmonic_rhs(fc, as_numpy=False)
    fv, info1 = lap.solve(b1, precond=precond, precision=precision, tol=tol, **kw)
    b2 = torch.empty((lap.n_free, 3 * P), dtype=torch.float64, device="cuda")
    L.check(lib.cn_deform_rhs(lap.handle, P, _p(fv), _p(fc), _p(xc), _p(b2), _stream()),
            "deform_rhs")
    xv, info2 = lap.solve(b2, precond=precond, precision=precision, tol=tol, **kw)
    free_of, con_of = lap.maps()
    fo = torch.as_tensor(free_of, device="cuda").long()
    co = torch.as_tensor(con_of, device="cuda").long()
    xa = torch.zeros((len(free_of), 3 * P), dtype=torch.float64, device="cuda")
    xa[fo >= 0] = xv[fo[fo >= 0]]
    xa[co >= 0] = xc[co[co >= 0]]
    return xv, xa, info1, info2


def bezier3(ctrl, t):
    """EndToEndExample.bezier3 on the GPU: ctrl (ncurves x 4 x 3), t (nt)
    -> ncurves x nt x 3 (cn_bezier3)."""
    torch = _torch()
    c = _as_tensor(np.asarray(ctrl, np.float64).reshape(-1, 12), torch.float64).contiguous()
    tv = _as_tensor(np.asarray(t, np.float64).reshape(-1), torch.float64).contiguous()
    out = torch.empty((c.shape[0], tv.shape[0], 3), dtype=torch.float64, device="cuda")
    L.check(L.load().cn_bezier3(c.shape[0], _p(c), tv.shape[0], _p(tv), _p(out), _stream()),
            "bezier3")
    return out.cpu().numpy()


def segment_gradients(rest_p, rest_q, posed_p, posed_q):
    """ScaledFrames.isolatedSegmentGradient per segment (cn_segment_gradients):
    n x 3 points each -> n x 3 x 3."""
    torch = _torch()
    a = [_as_tensor(np.asarray(v, np.float64).reshape(-1, 3), torch.float64).contiguous()
         for v in (rest_p, rest_q, posed_p, posed_q)]
    n = a[0].shape[0]
    F = torch.empty((n, 9), dtype=torch.float64, device="cuda")
    L.check(L.load().cn_segment_gradients(n, *[_p(v) for v in a], _p(F), _stream()),
            "segment_gradients")
    return F.cpu().numpy().reshape(n, 3, 3)


def deformation_gradients(rest, posed):
    """ScaledFrames.deformationGradient per frame pair: n x 3 x 3 each."""
    torch = _torch()
    r = _as_tensor(np.asarray(rest, np.float64).reshape(-1, 9), torch.float64).contiguous()
    p = _as_tensor(np.asarray(posed, np.float64).reshape(-1, 9), torch.float64).contiguous()
    F = torch.empty_like(r)
    L.check(L.load().cn_deformation_gradients(r.shape[0], _p(r), _p(p), _p(F), _stream()),
            "deformation_gradients")
    return F.cpu().numpy().reshape(-1, 3, 3)


class BatchSystem(_System):
    """Independent systems solved together (cn_system_batch): a
    block-diagonal copy of their A_ff.  Right-hand sides and solutions stack
    the members' rows in order; every (member, column) pair follows its own
    CG recurrence, so a member's rows of the solution solve its system."""

    def __init__(self, systems):
        lib = L.load()
        hs = (C.c_void_p * len(systems))(*[m.handle.value for m in systems])
        h = C.c_void_p(0)
        L.check(lib.cn_system_batch(len(systems), C.cast(hs, C.POINTER(C.c_void_p)), C.byref(h),
                                    _stream()), "system_batch")
        super().__init__(h.value)
        nseg = C.c_int32(0)
        sp = L.P_i32()
        L.check(lib.cn_system_segments(self._h, C.byref(nseg), C.byref(sp)), "system_segments")
        self.seg_ptr = np.array([sp[i] for i in range(nseg.value + 1)], np.int64)

    @property
    def nseg(self) -> int:
        return len(self.seg_ptr) - 1

    def harmonic_rhs(self, Xcs, as_numpy: bool = True):
        """Every member's HarmonicSolve right-hand side in one launch: Xcs is
        a list of the members' n_con_i x k constrained values (or those
        stacked); returns the stacked (sum of n_free) x k block."""
        torch = _torch()
        lib = L.load()
        x = _as_tensor(self.stack(Xcs) if isinstance(Xcs, (list, tuple)) else Xcs,
                       torch.float64)
        x = x.reshape(x.shape[0], -1).contiguous()
        k = x.shape[1]
        out = torch.empty((self.n_free, k), dtype=torch.float64, device="cuda")
        L.check(lib.cn_harmonic_rhs(self._h, k, _p(x), _p(out), _stream()), "harmonic_rhs")
        return out.cpu().numpy() if as_numpy else out

    def stack(self, blocks):
        """Concatenate per-member n_i x k blocks (numpy or device tensors)."""
        torch = _torch()
        if all(isinstance(b, torch.Tensor) for b in blocks):
            return torch.cat([b.reshape(b.shape[0], -1) for b in blocks], dim=0).contiguous()
        return np.concatenate([np.asarray(b).reshape(len(b), -1) for b in blocks], axis=0)

    def split(self, x):
        """The members' row blocks of a stacked block."""
        return [x[self.seg_ptr[i]:self.seg_ptr[i + 1]] for i in range(self.nseg)]


class HSC:
    """Hierarchical Schur preconditioner of a system's A_ff (cn_hsc_build):
    greedy independent-set elimination with exact Schur complements,
    sparsified fill, dense coarse inverse on the matrix cores."""

    def __init__(self, system: "_System", theta: float = 0.015, coarse_max: int = 512,
                 max_levels: int = 24, min_frac: float = 0.05, compensate: str = "lump"):
        lib = L.load()
        self.system = system
        o = L.HscOpts()
        lib.cn_hsc_default_opts(C.byref(o))
        o.theta, o.coarse_max, o.max_levels, o.min_frac = theta, coarse_max, max_levels, min_frac
        o.compensate = {"lump": L.CN_HSC_LUMP, "diag": L.CN_HSC_DIAG}[compensate]
        h = C.c_void_p(0)
        L.check(lib.cn_hsc_build(system.handle, C.byref(o), C.byref(h), _stream()), "hsc_build")
        self._adopt(h, o)

    def _adopt(self, h, o):
        self._h = h
        self.opts = o
        info = L.HscInfo()
        L.check(L.load().cn_hsc_get_info(h, C.byref(info)), "hsc_get_info")
        self.info = info

    @classmethod
    def from_import(cls, system: "_System", opts, levels, nlevels: int, n_coarse: int,
                    coarse_L: np.ndarray, coarse_inv: np.ndarray) -> "HSC":
        """A hierarchy rebuilt from persisted host arrays (cn_hsc_import)."""
        self = cls.__new__(cls)
        self.system = system
        h = C.c_void_p(0)
        L.check(L.load().cn_hsc_import(system.handle, C.byref(opts), nlevels, levels, n_coarse,
                                       coarse_L.ctypes.data if n_coarse else None,
                                       coarse_inv.ctypes.data if n_coarse else None,
                                       C.byref(h), _stream()), "hsc_import")
        self._adopt(h, opts)
        return self

    def tail_level(self, k: int, fp32: bool = True) -> int:
        """First level an apply of k columns runs in the fused tail launch
        (cn_hsc_tail_level), -1 when every level runs its own kernels."""
        return int(L.load().cn_hsc_tail_level(self._h, int(k), 1 if fp32 else 0))

    def save(self, path) -> None:
        """Write the hierarchy to a safetensors sidecar (formats.save_sidecar)."""
        from . import formats
        formats.save_sidecar(path, self.system, hsc=self)

    @staticmethod
    def load(system: "_System", path) -> "HSC":
        """The hierarchy of a sidecar written for `system` (formats.load_hsc)."""
        from . import formats
        return formats.load_hsc(path, system)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                L.load().cn_hsc_destroy(h)
            except Exception:
                pass
            self._h = C.c_void_p(0)

    @property
    def handle(self):
        return self._h

    def levels(self) -> list:
        """Host copies of every level: dict with n, F, C, dinv, R = (ptr, col,
        val) (columns are level rows of F vertices), P = (ptr, col, val)."""
        out = []
        for l in range(self.info.nlevels):
            v = L.HscLevel()
            L.check(L.load().cn_hsc_get_level(self._h, l, C.byref(v)), "hsc_get_level")
            rp = _np(v.d_r_ptr, v.nC + 1, np.int32)
            pp = _np(v.d_p_ptr, v.nF + 1, np.int32)
            out.append({"n": v.n, "F": _np(v.d_F, v.nF, np.int32), "C": _np(v.d_C, v.nC, np.int32),
                        "dinv": _np(v.d_dinv, v.nF, np.float64),
                        "R": (rp, _np(v.d_r_col, int(rp[-1]), np.int32),
                              _np(v.d_r_val, int(rp[-1]), np.float64)),
                        "P": (pp, _np(v.d_p_col, int(pp[-1]), np.int32),
                              _np(v.d_p_val, int(pp[-1]), np.float64))})
        return out

    def coarse(self):
        """(Cholesky factor, inverse) of the coarse matrix, dense FP64."""
        nc = self.info.n_coarse
        return (_np(self.info.d_coarse_L, nc * nc, np.float64).reshape(nc, nc),
                _np(self.info.d_coarse_inv, nc * nc, np.float64).reshape(nc, nc))

    def apply(self, r):
        torch = _torch()
        rt = _as_tensor(r, torch.float64)
        vec = rt.dim() == 1
        rt = rt.reshape(self.system.n_free, -1).contiguous()
        z = torch.empty_like(rt)
        L.check(L.load().cn_hsc_apply(self._h, rt.shape[1], _p(rt), _p(z), _stream()),
                "hsc_apply")
        z = z.cpu().numpy()
        return z.reshape(-1) if vec else z


class SGS:
    """Colour-parallel symmetric Gauss-Seidel on a system's A_ff
    (SlangCodegen/SgsColor.lean:87-130; cn_sgs_build): greedy colours,
    sweeps over colours 0..C-1 then C-1..0."""

    def __init__(self, system: "_System"):
        self.system = system
        h = C.c_void_p(0)
        L.check(L.load().cn_sgs_build(system.handle, C.byref(h), _stream()), "sgs_build")
        self._h = h
        info = L.SgsInfo()
        L.check(L.load().cn_sgs_get_info(h, C.byref(info)), "sgs_get_info")
        self.n, self.ncolors = info.n, info.ncolors
        self.color_ptr = np.array([info.h_color_ptr[c] for c in range(self.ncolors + 1)],
                                  np.int32)
        self.color_rows = _np(info.d_color_rows, self.n, np.int32)
        self.color = _np(info.d_color, self.n, np.int32)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                L.load().cn_sgs_destroy(h)
            except Exception:
                pass
            self._h = C.c_void_p(0)

    @property
    def handle(self):
        return self._h

    def sweep(self, b, x0, sweeps: int = 1, dtype=np.float64):
        """`sweeps` symmetric sweeps of A_ff x = b from x0 (n or n x k), in
        FP64 or FP32 (dtype).  Returns x on the host."""
        torch = _torch()
        tt = torch.float32 if dtype == np.float32 else torch.float64
        bt = _as_tensor(np.asarray(b, dtype=dtype), tt)
        vec = bt.dim() == 1
        bt = bt.reshape(self.n, -1).contiguous()
        x = _as_tensor(np.asarray(x0, dtype=dtype), tt).reshape(self.n, -1).contiguous().clone()
        fn = L.load().cn_sgs_sweep_f32 if dtype == np.float32 else L.load().cn_sgs_sweep_f64
        L.check(fn(self._h, bt.shape[1], _p(bt), _p(x), sweeps, _stream()), "sgs_sweep")
        x = x.cpu().numpy()
        return x.reshape(-1) if vec else x

    def apply(self, r):
        """z = M^-1 r: one symmetric sweep from zero (cn_sgs_apply)."""
        torch = _torch()
        rt = _as_tensor(r, torch.float64)
        vec = rt.dim() == 1
        rt = rt.reshape(self.n, -1).contiguous()
        z = torch.empty_like(rt)
        L.check(L.load().cn_sgs_apply(self._h, rt.shape[1], _p(rt), _p(z), _stream()),
                "sgs_apply")
        z = z.cpu().numpy()
        return z.reshape(-1) if vec else z


class Restriction:
    """Coarse-vertex restriction of a parent map (SlangCodegen/
    HierarchicalSparsify.lean:64-93; cn_restrict_build): coarse[c] = sum of
    fine[i] over parent[i] == c, in ascending i."""

    def __init__(self, parent, n_coarse: int):
        torch = _torch()
        p = _dev(np.asarray(parent, np.int32), torch.int32)
        self.n_fine, self.n_coarse = int(p.numel()), int(n_coarse)
        h = C.c_void_p(0)
        L.check(L.load().cn_restrict_build(self.n_fine, _p(p), self.n_coarse, C.byref(h),
                                           _stream()), "restrict_build")
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                L.load().cn_restrict_destroy(h)
            except Exception:
                pass
            self._h = C.c_void_p(0)

    def __call__(self, fine, dtype=np.float64):
        torch = _torch()
        tt = torch.float32 if dtype == np.float32 else torch.float64
        f = np.asarray(fine, dtype=dtype)
        ft = _as_tensor(f.reshape(self.n_fine, -1), tt).contiguous()
        k = ft.shape[1]
        out = torch.empty((self.n_coarse, k), dtype=tt, device="cuda")
        fn = L.load().cn_restrict_sum_f32 if dtype == np.float32 else L.load().cn_restrict_sum_f64
        L.check(fn(self._h, k, _p(ft), _p(out), _stream()), "restrict_sum")
        return out.cpu().numpy().reshape((self.n_coarse,) + f.shape[1:])


def dense_cholesky(A):
    """Lower Cholesky factor of a dense SPD matrix (cn_dense_cholesky_f64)."""
    torch = _torch()
    a = _as_tensor(np.asarray(A, dtype=np.float64), torch.float64).clone()
    n = a.shape[0]
    L.check(L.load().cn_dense_cholesky_f64(n, _p(a), n, _stream()), "dense_cholesky_f64")
    return a.cpu().numpy()


def dense_cholesky_solve(Lf, B):
    """(L L^T)^-1 B for a factor from dense_cholesky (cn_dense_cholesky_solve_f64)."""
    torch = _torch()
    lt = _as_tensor(np.asarray(Lf, dtype=np.float64), torch.float64)
    b = _as_tensor(np.asarray(B, dtype=np.float64), torch.float64)
    vec = b.dim() == 1
    b = b.reshape(lt.shape[0], -1).contiguous().clone()
    L.check(L.load().cn_dense_cholesky_solve_f64(lt.shape[0], b.shape[1], _p(lt), lt.shape[0],
                                                 _p(b), b.shape[1], _stream()),
            "dense_cholesky_solve_f64")
    out = b.cpu().numpy()
    return out.reshape(-1) if vec else out


class IC0:
    """No-fill incomplete Cholesky factor of a system's A_ff."""

    def __init__(self, system: _System, ordering: str = "block", allow_shift: bool = True,
                 block_size: int = 16):
        lib = L.load()
        self.system = system
        h = C.c_void_p(0)
        o = L.Ic0Opts()
        o.ordering = {"color": L.CN_ORDER_COLOR, "natural": L.CN_ORDER_NATURAL,
                      "block": L.CN_ORDER_BLOCK}[ordering]
        o.block_size = block_size
        o.allow_shift = int(allow_shift)
        L.check(lib.cn_ic0_factor_ex(system.handle, C.byref(o), C.byref(h), _stream()),
                "ic0_factor")
        self._h = h
        info = L.Ic0Info()
        L.check(lib.cn_ic0_get_info(h, C.byref(info)), "ic0_get_info")
        self.info = info

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                L.load().cn_ic0_destroy(h)
            except Exception:
                pass
            self._h = C.c_void_p(0)

    @property
    def handle(self):
        return self._h

    def perm(self) -> np.ndarray:
        return _np(self.info.d_perm, self.info.n, np.int32)

    def permuted_matrix(self) -> CsrMatrix:
        i = self.info
        n = i.n
        ptr = _np(i.d_a_ptr, n + 1, np.int32)
        return CsrMatrix(n, n, ptr, _np(i.d_a_col, int(ptr[-1]), np.int32),
                         _np(i.d_a_val, int(ptr[-1]), np.float64))

    def factor(self):
        """(l_ptr, l_col, l_val, diag) of the strictly lower part + diagonal."""
        i = self.info
        n = i.n
        ptr = _np(i.d_l_ptr, n + 1, np.int32)
        nnz = int(ptr[-1])
        return ptr, _np(i.d_l_col, nnz, np.int32), _np(i.d_l_val, nnz, np.float64), \
            _np(i.d_l_diag, n, np.float64)

    def apply(self, r):
        torch = _torch()
        rt = _as_tensor(r, torch.float64)
        vec = rt.dim() == 1
        rt = rt.reshape(self.info.n, -1).contiguous()
        z = torch.empty_like(rt)
        L.check(L.load().cn_ic0_apply(self._h, rt.shape[1], _p(rt), _p(z), _stream()), "ic0_apply")
        z = z.cpu().numpy()
        return z.reshape(-1) if vec else z


# ------------------------------------------------------------- spec mirrors --
def cg_with_guess(A: CsrMatrix, b, x0, maxIter: int, tol: float):
    """SparseLinAlg.cgWithGuess: Jacobi PCG from x0, stop when ||r|| < tol
    (absolute, as the spec).  FP64 on the GPU."""
    b = np.asarray(b, dtype=np.float64).reshape(-1)
    nb = float(np.linalg.norm(b))
    sysm = CsrSystem(A)
    if nb == 0.0:
        return np.zeros_like(b)
    x, info = sysm.solve(b, precond="jacobi", precision="fp64_plain", tol=tol / nb, x0=x0,
                         max_iter=maxIter)
    return x.cpu().numpy()


def cg(A: CsrMatrix, b, maxIter: int, tol: float):
    """SparseLinAlg.cg (SparseLinAlg.lean:104-130)."""
    return cg_with_guess(A, b, None, maxIter, tol)


def harmonic_solve(lap: CutLaplacian, Xc, **solve_kw):
    """HarmonicSolve.solveMulti on the sparse path: returns the nv x k block
    with free rows solved and sample rows 0 (the caller overlays them,
    HarmonicSolve.lean:14-16)."""
    b = lap.harmonic_rhs(Xc, as_numpy=False)
    x, info = lap.solve(b, **solve_kw)
    free_of, _ = lap.maps()
    out = np.zeros((lap.info.nv, b.shape[1]))
    xf = x.cpu().numpy()
    sel = free_of >= 0
    out[sel] = xf[free_of[sel]]
    return out, info
