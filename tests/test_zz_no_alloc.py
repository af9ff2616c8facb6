"""cn_hsc_apply_noalloc: the HSC apply with its workspace in the library's
fixed per-GPU scratch array instead of an allocation.  Checked through the
C ABI against cn_hsc_apply (the same kernels: the same bits) on two calls
in a row, the second with fewer columns, with the device's free memory
unchanged across them; an apply whose workspace exceeds the scratch is
refused before any launch."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _apply(lib, fn, H, r):
    import torch
    rt = torch.as_tensor(np.ascontiguousarray(r), device="cuda")
    zt = torch.empty_like(rt)
    st = fn(H.handle, rt.shape[1], C.c_void_p(rt.data_ptr()), C.c_void_p(zt.data_ptr()),
            C.c_void_p(torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    return st, zt.cpu().numpy()


@pytest.mark.parametrize("name,ks", [("icosphere6", (48, 5)), ("cylinder64", (300, 3))])
def test_noalloc_apply_equals_apply(gpu, name, ks):
    import torch
    from tool_godot_curvenet_amd import HSC, CutLaplacian, _lib, meshgen
    lib = _lib.load()
    lap = CutLaplacian.from_mesh(meshgen.config_mesh(name))
    H = HSC(lap)
    rng = np.random.default_rng(5)
    rs = [rng.normal(size=(lap.n_free, k)) for k in ks]
    # the allocating entry first (its workspace, the kernels' code objects)
    want = []
    for r in rs:
        st, z = _apply(lib, lib.cn_hsc_apply, H, r)
        assert st == 0
        want.append(z)
    rt = [torch.as_tensor(r, device="cuda") for r in rs]
    zt = [torch.empty_like(t) for t in rt]
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    for t, z in zip(rt, zt):
        st = lib.cn_hsc_apply_noalloc(H.handle, t.shape[1], C.c_void_p(t.data_ptr()),
                                      C.c_void_p(z.data_ptr()),
                                      C.c_void_p(torch.cuda.current_stream().cuda_stream))
        assert st == 0, lib.cn_last_error()
    torch.cuda.synchronize()
    assert torch.cuda.mem_get_info()[0] == free0
    for z, w in zip(zt, want):
        assert np.array_equal(z.cpu().numpy(), w)


def test_noalloc_rejects_oversize(gpu):
    import torch
    from tool_godot_curvenet_amd import HSC, CutLaplacian, _lib, meshgen
    lib = _lib.load()
    lap = CutLaplacian.from_mesh(meshgen.config_mesh("cylinder64"))
    H = HSC(lap)
    nc = H.info.n_coarse
    k = (16 << 20) // (8 * nc) + 8              # nc x k x 8 bytes > the 16 MiB scratch
    r = torch.zeros((lap.n_free, k), dtype=torch.float64, device="cuda")
    z = torch.full_like(r, 7.0)
    st = lib.cn_hsc_apply_noalloc(H.handle, k, C.c_void_p(r.data_ptr()),
                                  C.c_void_p(z.data_ptr()), None)
    assert st == _lib.CN_BAD_ARG
    assert b"scratch" in lib.cn_last_error()
    torch.cuda.synchronize()
    assert bool((z == 7.0).all())               # nothing was launched
