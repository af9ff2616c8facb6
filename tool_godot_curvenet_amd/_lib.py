"""ctypes binding of ``libcurvenet_la.so`` (the C ABI in include/curvenet_la.h).

The library is loaded from the package directory.  A missing library is an
error: there is no fallback implementation.
"""
from __future__ import annotations

import ctypes as C
import pathlib
import re

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "libcurvenet_la.so"
HEADER = HERE.parent / "include" / "curvenet_la.h"

CN_OK, CN_MAXIT, CN_BREAKDOWN, CN_NOT_SPD = 0, 1, 2, 3
CN_BAD_ARG, CN_HIP_ERROR, CN_NO_DEVICE, CN_OOM, CN_NOT_MANIFOLD = -1, -2, -3, -4, -5
STATUS_NAMES = {0: "OK", 1: "MAXIT", 2: "BREAKDOWN", 3: "NOT_SPD", -1: "BAD_ARG",
                -2: "HIP_ERROR", -3: "NO_DEVICE", -4: "OOM", -5: "NOT_MANIFOLD"}

CN_ORDER_COLOR, CN_ORDER_NATURAL, CN_ORDER_BLOCK = 0, 1, 2
CN_PRECOND_JACOBI, CN_PRECOND_IC0, CN_PRECOND_HSC, CN_PRECOND_SGS = 0, 1, 2, 3
CN_HSC_LUMP, CN_HSC_DIAG = 0, 1
CN_PREC_FP64_PLAIN, CN_PREC_FP64, CN_PREC_MIXED = 0, 1, 2

i32 = C.c_int32
i64 = C.c_int64
f64 = C.c_double
vp = C.c_void_p
P_i32 = C.POINTER(C.c_int32)


class SystemInfo(C.Structure):
    _fields_ = [("nv", i32), ("nf", i32), ("nh", i32), ("n_free", i32), ("n_con", i32),
                ("nnz_ff", i64), ("nnz_fc", i64),
                ("d_ff_ptr", vp), ("d_ff_col", vp), ("d_ff_val", vp), ("d_ff_val32", vp),
                ("d_fc_ptr", vp), ("d_fc_col", vp), ("d_fc_val", vp),
                ("d_free_of", vp), ("d_con_of", vp), ("d_free_vertex", vp),
                ("d_con_vertex", vp), ("d_he_target", vp), ("d_he_twin", vp),
                ("d_he_next", vp), ("d_he_face", vp)]


CN_LAP_COT, CN_LAP_ROBUST = 0, 1


class SystemOpts(C.Structure):
    _fields_ = [("laplacian", i32), ("delta", f64)]


class Ic0Opts(C.Structure):
    _fields_ = [("ordering", i32), ("block_size", i32), ("allow_shift", i32)]


class Ic0Info(C.Structure):
    _fields_ = [("n", i32), ("nlevels", i32), ("ordering", i32), ("nblocks", i32),
                ("max_block", i32), ("nnz_l", i64),
                ("shift", f64), ("d_perm", vp), ("d_a_ptr", vp), ("d_a_col", vp),
                ("d_a_val", vp), ("d_l_ptr", vp), ("d_l_col", vp), ("d_l_val", vp),
                ("d_l_diag", vp)]


class PcgOpts(C.Structure):
    _fields_ = [("precond", i32), ("precision", i32), ("tol_rel", f64), ("max_iter", i32),
                ("inner_tol", f64), ("max_refine", i32), ("check_every", i32),
                ("warm_start", i32), ("graph", i32)]


class Tune(C.Structure):
    _fields_ = [("spmm_form", i32), ("spmm_gens", i32), ("level_cap", i32), ("tail_rows", i32),
                ("chain_levels", i32), ("coarse_prec", i32), ("down_cap", i32)]


class HscOpts(C.Structure):
    _fields_ = [("theta", f64), ("coarse_max", i32), ("max_levels", i32), ("min_frac", f64),
                ("compensate", i32)]


class HscInfo(C.Structure):
    _fields_ = [("nlevels", i32), ("n_coarse", i32), ("nnz", i64), ("d_coarse_L", vp),
                ("d_coarse_inv", vp)]


class HscLevel(C.Structure):
    _fields_ = [("n", i32), ("nF", i32), ("nC", i32), ("d_F", vp), ("d_C", vp), ("d_dinv", vp),
                ("d_r_ptr", vp), ("d_r_col", vp), ("d_r_val", vp), ("d_p_ptr", vp),
                ("d_p_col", vp), ("d_p_val", vp)]


class HscLevelIn(C.Structure):
    _fields_ = [("n", i32), ("nF", i32), ("nC", i32), ("F", vp), ("C", vp), ("dinv", vp),
                ("r_ptr", vp), ("r_col", vp), ("r_val", vp), ("p_ptr", vp), ("p_col", vp),
                ("p_val", vp)]


class StTensor(C.Structure):
    _fields_ = [("name", C.c_char_p), ("dtype", i32), ("ndim", i32),
                ("shape", C.POINTER(i64)), ("data", vp), ("nbytes", i64)]


class SgsInfo(C.Structure):
    _fields_ = [("n", i32), ("ncolors", i32), ("h_color_ptr", P_i32), ("d_color_rows", vp),
                ("d_color", vp)]


class SolveInfo(C.Structure):
    _fields_ = [("status", i32), ("iters", i32), ("refinements", i32), ("relres_max", f64)]


_SIGS = {
    "cn_last_error": (C.c_char_p, []),
    "cn_version": (i32, []),
    "cn_device_check": (i32, []),
    "cn_halfedge_from_triangles": (i32, [i32, i32, vp, vp, vp, vp, vp, P_i32, vp]),
    "cn_cut_subdivide_edge": (i32, [P_i32, P_i32, P_i32, i32, vp, vp, vp, vp]),
    "cn_cut_split_face": (i32, [P_i32, P_i32, P_i32, i32, i32, vp, vp, vp, vp]),
    "cn_cut_insert_crack": (i32, [P_i32, P_i32, P_i32, i32, vp, vp, vp, vp]),
    "cn_cut_seams": (i32, [i32, i32, vp, i32, vp, vp, vp, vp, P_i32, vp]),
    "cn_project_to_vertices": (i32, [i32, vp, i32, vp, vp, vp]),
    "cn_promote_vertex_samples": (i32, [i32, i32, vp, vp, vp, vp, vp]),
    "cn_system_create": (i32, [i32, i32, vp, vp, vp, C.POINTER(vp), vp]),
    "cn_system_default_opts": (None, [C.POINTER(SystemOpts)]),
    "cn_system_create_ex": (i32, [i32, i32, vp, vp, vp, C.POINTER(SystemOpts), C.POINTER(vp),
                                  vp]),
    "cn_system_from_csr": (i32, [i32, vp, vp, vp, C.POINTER(vp), vp]),
    "cn_system_get_info": (i32, [vp, C.POINTER(SystemInfo)]),
    "cn_system_set_positions": (i32, [vp, vp, vp]),
    "cn_system_destroy": (i32, [vp]),
    "cn_system_batch": (i32, [i32, C.POINTER(vp), C.POINTER(vp), vp]),
    "cn_system_segments": (i32, [vp, P_i32, C.POINTER(P_i32)]),
    "cn_system_set_tune": (i32, [vp, C.POINTER(Tune)]),
    "cn_system_get_tune": (i32, [vp, C.POINTER(Tune)]),
    "cn_harmonic_rhs": (i32, [vp, i32, vp, vp, vp]),
    "cn_spmv_f64": (i32, [i32, vp, vp, vp, vp, vp, vp]),
    "cn_spmv_f32": (i32, [i32, vp, vp, vp, vp, vp, vp]),
    "cn_spmm_f64": (i32, [i32, vp, vp, vp, i32, vp, vp, vp]),
    "cn_spmm_f32": (i32, [i32, vp, vp, vp, i32, vp, vp, vp]),
    "cn_ic0_factor": (i32, [vp, i32, i32, C.POINTER(vp), vp]),
    "cn_ic0_factor_ex": (i32, [vp, C.POINTER(Ic0Opts), C.POINTER(vp), vp]),
    "cn_ic0_get_info": (i32, [vp, C.POINTER(Ic0Info)]),
    "cn_ic0_apply": (i32, [vp, i32, vp, vp, vp]),
    "cn_ic0_destroy": (i32, [vp]),
    "cn_pcg_default_opts": (None, [C.POINTER(PcgOpts)]),
    "cn_pcg_solve": (i32, [vp, vp, C.POINTER(PcgOpts), i32, vp, vp, C.POINTER(SolveInfo), vp]),
    "cn_adjoint_solve": (i32, [vp, vp, C.POINTER(PcgOpts), i32, vp, vp, vp, vp, vp, vp,
                               C.POINTER(SolveInfo), vp]),
    "cn_adjoint_solve_hsc": (i32, [vp, vp, C.POINTER(PcgOpts), i32, vp, vp, vp, vp, vp, vp,
                                   C.POINTER(SolveInfo), vp]),
    "cn_pcg_solve_hsc": (i32, [vp, vp, C.POINTER(PcgOpts), i32, vp, vp, C.POINTER(SolveInfo),
                               vp]),
    "cn_dense_cholesky_f64": (i32, [i32, vp, i32, vp]),
    "cn_dense_cholesky_solve_f64": (i32, [i32, i32, vp, i32, vp, i32, vp]),
    "cn_hsc_default_opts": (None, [C.POINTER(HscOpts)]),
    "cn_hsc_build": (i32, [vp, C.POINTER(HscOpts), C.POINTER(vp), vp]),
    "cn_hsc_get_info": (i32, [vp, C.POINTER(HscInfo)]),
    "cn_hsc_get_level": (i32, [vp, i32, C.POINTER(HscLevel)]),
    "cn_hsc_apply": (i32, [vp, i32, vp, vp, vp]),
    "cn_hsc_apply_noalloc": (i32, [vp, i32, vp, vp, vp]),
    "cn_hsc_destroy": (i32, [vp]),
    "cn_sgs_build": (i32, [vp, C.POINTER(vp), vp]),
    "cn_sgs_get_info": (i32, [vp, C.POINTER(SgsInfo)]),
    "cn_sgs_sweep_f64": (i32, [vp, i32, vp, vp, i32, vp]),
    "cn_sgs_sweep_f32": (i32, [vp, i32, vp, vp, i32, vp]),
    "cn_sgs_apply": (i32, [vp, i32, vp, vp, vp]),
    "cn_sgs_destroy": (i32, [vp]),
    "cn_pcg_solve_sgs": (i32, [vp, vp, C.POINTER(PcgOpts), i32, vp, vp, C.POINTER(SolveInfo),
                               vp]),
    "cn_restrict_build": (i32, [i32, vp, i32, C.POINTER(vp), vp]),
    "cn_restrict_sum_f64": (i32, [vp, i32, vp, vp, vp]),
    "cn_restrict_sum_f32": (i32, [vp, i32, vp, vp, vp]),
    "cn_restrict_destroy": (i32, [vp]),
    "cn_deform_rhs": (i32, [vp, i32, vp, vp, vp, vp, vp]),
    "cn_bezier3": (i32, [i32, vp, i32, vp, vp, vp]),
    "cn_segment_gradients": (i32, [i32, vp, vp, vp, vp, vp, vp]),
    "cn_deformation_gradients": (i32, [i32, vp, vp, vp, vp]),
    "cn_hsc_tail_level": (i32, [vp, i32, i32]),
    "cn_hsc_import": (i32, [vp, C.POINTER(HscOpts), i32, C.POINTER(HscLevelIn), i32, vp, vp,
                            C.POINTER(vp), vp]),
    "cn_st_parse": (i32, [vp, i64, C.POINTER(vp)]),
    "cn_st_read_file": (i32, [C.c_char_p, C.POINTER(vp)]),
    "cn_st_count": (i32, [vp]),
    "cn_st_tensor_at": (i32, [vp, i32, C.POINTER(StTensor)]),
    "cn_st_find": (i32, [vp, C.c_char_p]),
    "cn_st_metadata_count": (i32, [vp]),
    "cn_st_metadata_at": (i32, [vp, i32, C.POINTER(C.c_char_p), C.POINTER(C.c_char_p)]),
    "cn_st_metadata_get": (C.c_char_p, [vp, C.c_char_p]),
    "cn_st_free": (None, [vp]),
    "cn_st_encode": (i32, [i32, C.POINTER(StTensor), i32, C.POINTER(C.c_char_p),
                           C.POINTER(C.c_char_p), vp, i64, C.POINTER(i64)]),
    "cn_st_write_file": (i32, [C.c_char_p, i32, C.POINTER(StTensor), i32,
                               C.POINTER(C.c_char_p), C.POINTER(C.c_char_p)]),
    "cn_lbs_matvec_f64": (i32, [i32, i32, i32, vp, vp, vp, vp, vp, vp, vp]),
    "cn_lbs_matvec_f32": (i32, [i32, i32, i32, vp, vp, vp, vp, vp, vp, vp]),
    "cn_smooth_weights_f64": (i32, [i32, i32, vp, vp, vp, vp, i32, f64, vp]),
    "cn_sparsify_topk_f64": (i32, [i32, i32, vp, i32, vp, vp, vp]),
    "cn_bind_rhs": (i32, [vp, i32, i32, vp, vp]),
    "cn_bind_weights": (i32, [vp, i32, i32, vp, vp, i64, vp]),
    "cn_profile_enable": (None, [i32]),
    "cn_profile_query": (i32, [C.c_char_p, C.POINTER(f64)]),
    "cn_profile_only": (None, [C.c_char_p]),
}

_lib = None


def declared_symbols() -> list[str]:
    """Function names declared in include/curvenet_la.h."""
    txt = HEADER.read_text()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(cn_[a-z0-9_]+)\s*\(", txt)))


def load() -> C.CDLL:
    """Load the library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: build it with "
                           "`python -m tool_godot_curvenet_amd.build` (no CPU fallback exists)")
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class CurvenetError(RuntimeError):
    def __init__(self, status: int, where: str):
        msg = load().cn_last_error()
        msg = msg.decode() if msg else ""
        super().__init__(f"{where}: {STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


def check(status: int, where: str, ok=(CN_OK,)) -> int:
    if status not in ok:
        raise CurvenetError(status, where)
    return status


# --------------------------------------------------------- raw device I/O --
_hip = None


def _hiprt():
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipMemcpy.restype = C.c_int
        _hip.hipMemcpy.argtypes = [vp, vp, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.restype = C.c_int
    return _hip


def d2h(ptr: int, nbytes: int) -> bytes:
    """Copy nbytes from a device pointer to host (synchronous)."""
    buf = C.create_string_buffer(max(nbytes, 1))
    if nbytes:
        r = _hiprt().hipMemcpy(C.cast(buf, vp), vp(ptr), nbytes, 2)  # DeviceToHost
        if r != 0:
            raise RuntimeError(f"hipMemcpy D2H failed ({r})")
    return buf.raw[:nbytes]
