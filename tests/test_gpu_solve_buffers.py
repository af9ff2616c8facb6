"""The solver's output buffer: x accumulates in the caller's d_x, and a call
with d_x == d_b (in place) solves the same system as one with separate
buffers; a warm start from the converged x stops at once."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("k", [3, 48])
def test_in_place_solve_matches(gpu, k):
    import torch
    from tool_godot_curvenet_amd import HSC, CutLaplacian, _lib, meshgen
    from tool_godot_curvenet_amd import meshgen as G
    lap = CutLaplacian.from_mesh(meshgen.config_mesh("cylinder64"))
    H = HSC(lap)
    lib = _lib.load()
    B = torch.as_tensor(lap.harmonic_rhs(G.seeded_boundary(lap.n_con, k, seed=5)),
                        device="cuda").contiguous()
    opts = _lib.PcgOpts()
    lib.cn_pcg_default_opts(opts)
    opts.precision = _lib.CN_PREC_MIXED
    opts.tol_rel = 1e-9
    sp = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    info = _lib.SolveInfo()
    X = torch.full_like(B, 7.0)                 # overwritten: no warm start
    _lib.check(lib.cn_pcg_solve_hsc(lap.handle, H.handle, C.byref(opts), k,
                                    C.c_void_p(B.data_ptr()), C.c_void_p(X.data_ptr()),
                                    C.byref(info), sp), "pcg_solve_hsc")
    it1 = info.iters
    Y = B.clone()                               # in place: b in, x out
    _lib.check(lib.cn_pcg_solve_hsc(lap.handle, H.handle, C.byref(opts), k,
                                    C.c_void_p(Y.data_ptr()), C.c_void_p(Y.data_ptr()),
                                    C.byref(info), sp), "pcg_solve_hsc")
    assert info.iters == it1 and info.status == 0
    assert torch.equal(X, Y)
    # warm start from the solution: the first FP64 residual already passes
    opts.warm_start = 1
    Z = X.clone()
    _lib.check(lib.cn_pcg_solve_hsc(lap.handle, H.handle, C.byref(opts), k,
                                    C.c_void_p(B.data_ptr()), C.c_void_p(Z.data_ptr()),
                                    C.byref(info), sp), "pcg_solve_hsc")
    assert info.iters == 0 and info.status == 0
    assert torch.equal(Z, X)
