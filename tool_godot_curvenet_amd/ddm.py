"""Direct Delta Mush on the GPU: the runtime skinning matvec and the bind-time
weight pipeline (DirectDeltaMush.lean, DirectDeltaMushBind.lean,
SlangCodegen/DirectDeltaMush.lean), over the C ABI of ``include/curvenet_la.h``.

Arrays come in as numpy arrays or torch tensors and go out as device
tensors; nothing here computes on the host except the mesh adjacency
(connectivity bookkeeping before any weight exists).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .curvenet import HSC, CutLaplacian, _as_tensor, _p, _stream, _torch


def vertex_adjacency(nv: int, tris) -> tuple[np.ndarray, np.ndarray]:
    """Vertex neighbour lists of a triangle mesh as CSR (ptr[nv + 1], adj):
    the vertices sharing an edge with v, ascending, each once -- the
    `Adjacency` (Common.lean:23-25) that smoothWeights reads."""
    t = np.asarray(tris, np.int64).reshape(-1, 3)
    a = np.concatenate([t[:, 0], t[:, 1], t[:, 2], t[:, 1], t[:, 2], t[:, 0]])
    b = np.concatenate([t[:, 1], t[:, 2], t[:, 0], t[:, 0], t[:, 1], t[:, 2]])
    key = np.unique(a * nv + b)
    src, dst = key // nv, key % nv
    ptr = np.zeros(nv + 1, np.int32)
    np.add.at(ptr, src + 1, 1)
    return np.cumsum(ptr).astype(np.int32), dst.astype(np.int32)


def lbs_matvec(transforms, start, idx, w, rest, precision: str = "f64"):
    """DirectDeltaMush.lbsMatvec (DirectDeltaMush.lean:85-98) for every vertex
    and pose: transforms (P x nt x 16 or nt x 16, 4x4 row-major), influences
    CSR (start, idx, w), rest (nv x 3).  Returns P x nv x 3 (nv x 3 for one
    pose).  precision "f64" keeps the spec's operation order; "f32" is the
    shader's float arithmetic."""
    torch = _torch()
    dt = torch.float64 if precision == "f64" else torch.float32
    if precision not in ("f64", "f32"):
        raise ValueError("precision must be 'f64' or 'f32'")
    T = _as_tensor(transforms, dt)
    single = T.dim() == 2
    if single:
        T = T.unsqueeze(0)
    if T.dim() != 3 or T.shape[2] != 16:
        raise ValueError("transforms must be nt x 16 or P x nt x 16")
    T = T.contiguous()
    P, nt = T.shape[0], T.shape[1]
    r = _as_tensor(rest, dt).reshape(-1, 3).contiguous()
    nv = r.shape[0]
    st = _as_tensor(start, torch.int32).reshape(-1).contiguous()
    ix = _as_tensor(idx, torch.int32).reshape(-1).contiguous()
    wt = _as_tensor(w, dt).reshape(-1).contiguous()
    # the kernel reads idx[q], w[q] for q in [start[v], start[v+1]): the CSR
    # is checked on the host before any launch
    from .formats import check_influences
    check_influences(nv, st.cpu().numpy(), ix, wt)
    out = torch.empty((P, nv, 3), dtype=dt, device="cuda")
    fn = L.load().cn_lbs_matvec_f64 if precision == "f64" else L.load().cn_lbs_matvec_f32
    L.check(fn(nv, nt, P, _p(r), _p(T), _p(st), _p(ix), _p(wt), _p(out), _stream()),
            "lbs_matvec")
    return out[0] if single else out


def smooth_weights(W, adj_ptr, adj, nu: int, omega: float):
    """DirectDeltaMushBind.smoothWeights (DirectDeltaMushBind.lean:76-97):
    nu passes of W <- (1 - omega) W + omega (neighbour mean), adjacency CSR
    in the caller's neighbour order.  Returns a new n x nh device tensor."""
    torch = _torch()
    Wt = _as_tensor(W, torch.float64)
    if Wt.dim() != 2:
        raise ValueError("W must be n x nh")
    Wt = Wt.clone()
    n, nh = Wt.shape
    ap = _as_tensor(adj_ptr, torch.int32).reshape(-1).contiguous()
    aj = _as_tensor(adj, torch.int32).reshape(-1).contiguous()
    if ap.numel() != n + 1:
        raise ValueError("adj_ptr must have n + 1 entries")
    tmp = torch.empty_like(Wt)
    L.check(L.load().cn_smooth_weights_f64(n, nh, _p(ap), _p(aj), _p(Wt), _p(tmp), int(nu),
                                           C.c_double(omega), _stream()), "smooth_weights")
    return Wt


def sparsify_topk(W, K: int):
    """DirectDeltaMushBind.sparsifyTopK (DirectDeltaMushBind.lean:107-136):
    (idx, w), each n x min(K, nh), in the spec's selection order and
    renormalised to sum 1 where the kept sum is nonzero."""
    torch = _torch()
    Wt = _as_tensor(W, torch.float64)
    n, nh = Wt.shape
    take = min(int(K), nh)
    idx = torch.empty((n, take), dtype=torch.int32, device="cuda")
    w = torch.empty((n, take), dtype=torch.float64, device="cuda")
    L.check(L.load().cn_sparsify_topk_f64(n, nh, _p(Wt), int(K), _p(idx), _p(w), _stream()),
            "sparsify_topk")
    return idx, w


def partition_of_unity(W, eps: float) -> bool:
    """DirectDeltaMushBind.partitionOfUnity (DirectDeltaMushBind.lean:61-65):
    every row sum within eps of 1 (rows summed on the device)."""
    torch = _torch()
    Wt = _as_tensor(W, torch.float64)
    return bool(((Wt.sum(dim=1) - 1.0).abs() < eps).all().item())


def bind_harvest(lap: CutLaplacian, hsc: "HSC | None" = None, block: int = 768,
                 precond: str = "hsc", precision: str = "mixed", tol: float = 1e-9):
    """Unit-perturbation harmonic weights of every handle (constrained
    column): W (nv x n_con) with W[v, h] the harmonic solution for e_h at
    free vertices and the unit rows at sample vertices (DirectDeltaMushBind
    step 1, the sparse multi-RHS harvest of SparseBindHarvest.lean:3-5).
    Handles are solved `block` columns at a time.  Returns (W, infos)."""
    torch = _torch()
    lib = L.load()
    H = lap.n_con
    nv = lap.info.nv
    if precond == "hsc" and hsc is None:
        hsc = HSC(lap)
    kw = {"hsc": hsc} if precond == "hsc" else {}
    W = torch.empty((nv, H), dtype=torch.float64, device="cuda")
    infos = []
    for h0 in range(0, H, block):
        nb = min(block, H - h0)
        rhs = torch.empty((lap.n_free, nb), dtype=torch.float64, device="cuda")
        L.check(lib.cn_bind_rhs(lap.handle, h0, nb, _p(rhs), _stream()), "bind_rhs")
        x, info = lap.solve(rhs, precond=precond, precision=precision, tol=tol, **kw)
        L.check(lib.cn_bind_weights(lap.handle, h0, nb, _p(x), _p(W), H, _stream()),
                "bind_weights")
        infos.append(info)
    return W, infos


def bind(lap: CutLaplacian, tris, nu: int = 2, omega: float = 0.5, K: int = 8, **harvest_kw):
    """The bind pipeline (DirectDeltaMushBind.lean:9-27): harvest, smooth over
    the mesh's vertex adjacency, sparsify to K influences.  Returns the CSR
    influences (start, idx, w) of lbs_matvec as device tensors, and the
    harvest's solve infos."""
    torch = _torch()
    W, infos = bind_harvest(lap, **harvest_kw)
    ap, aj = vertex_adjacency(lap.info.nv, tris)
    Ws = smooth_weights(W, ap, aj, nu, omega)
    del W
    idx, w = sparsify_topk(Ws, K)
    n, take = idx.shape
    start = torch.arange(0, (n + 1) * take, take, dtype=torch.int32, device="cuda")
    return start, idx.reshape(-1), w.reshape(-1), infos
