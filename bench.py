#!/usr/bin/env python3
"""Benchmark: solves/s at 1e-9 relative residual on the ~100k-vertex cut
Laplacian (capsule union, 16 cut curves), 256 poses x 3 RHS per step,
mixed-precision PCG with all 768 columns in lockstep (SpMM), preconditioned
by the hierarchical Schur preconditioner (or IC(0) with --precond ic0).

One step = build the harmonic RHS of all 256 poses from their seeded
constraint displacements (HarmonicSolve RHS, FP64) + solve them to
||b - Ax|| <= 1e-9 ||b|| per column (FP64 recomputed residual).  Inputs are
resident in HBM before the timed region; the preconditioner is built once
per mesh (bind time) outside it.

Prints one JSON line (the contract) with `roofline` for the dominant kernel
of the inner loop (the kernel with the most GPU time, timed with HIP events
around every launch inside the timed region) and `cpu_baseline` (the
oracle's FP64 IC(0)-PCG on this host's cores).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
WORKLOAD = "capsules100k"
TRAFFIC_FILE = os.path.join(ROOT, "profiles", f"traffic_{WORKLOAD}.json")
# kernels whose HBM traffic is reported (name prefixes in the counter files)
TRAFFIC_KERNELS = {
    "spmm_dir": "void cn::kl_spmm<float, 6, 1>",
    "spmm_upd": "void cn::kl_spmm<float, 6, 2>",
    "spmm": "void cn::kl_spmm<float, 6, 0>",
    "spmm_row": "void cn::kw_spmm_dot<float, 4, 64>",
    "fwd": "void cn::kb_fwd<float, 4, 16, true>",
    "bwd": "void cn::kb_bwd<float, 4, 16>",
    "hsc_down": "void cn::kw_hsc_chain_rec<float, 4, 64>",
    "hsc_up": "void cn::kw_hsc_up<float, 4, 64>",
    "pd_update": "void cn::kw_pd_update<float, 4, 64>",
    "dot": "void cn::kw_dot<float, 4, 64>",
    # the x update and FP64 residual of a replacement (KL_RES)
    "resid": "void cn::kl_spmm<double, 6, 3>",
    "hsc_tail": "void cn::k_hsc_tail<float, 4>",
    # one coarse product: the MFMA kernel and the sum of its slices, on the
    # FP32 rounding of the coarse inverse (FP32 MFMA) or on the FP64 inverse
    "hsc_coarse": ("void cn::k_coarse_mm<float, float, 128, 128, 4, 2>",
                   "void cn::k_coarse_sum<float, float>"),
    "hsc_coarse_f64": ("void cn::k_coarse_mm<float, double, 128, 128, 4, 2>",
                       "void cn::k_coarse_sum<float, double>"),
    # on the inverse split into two bfloat16 matrices (coarse_prec 3)
    "hsc_coarse_bf3": ("void cn::k_coarse_mm_bf3<float, 128, 128, 4, 2>",
                       "void cn::k_coarse_sum<float, float>"),
}
FP64_MFMA_PEAK_TFLOPS = 78.6
FP32_MFMA_PEAK_TFLOPS = 157.3
BF16_MFMA_PEAK_TFLOPS = 2500.0
# HSC hierarchy per workload: (drop threshold, coarse rows at most, levels
# at most).  With 768 columns the dense coarse product runs on the FP64
# matrix cores at about 0.65 of peak, so a small threshold and a large coarse
# block pay: on the capsule union theta 0.015 / 512 rows takes 24
# iterations, 0.002 / 4096 rows 10.
# On the level-9 icosphere (k = 48) the 24-level cap binds first, so it gets
# more levels: theta 0.015 / 512 / 24 takes 29 + 36 iterations (76 poses/s),
# 0.002 / 8192 / 48 14 + 18 (117 poses/s).  With the FP32 coarse product the
# larger blocks pay: 0.002 at 8192 / 12288 / 16384 rows 130.4 / 133.9 / 136.5
# poses/s, 0.0015 at 12288 / 16384 137.3 / 132.2 (12 + 16 iterations), 0.001
# at 12288 135.6.
# The 64-mesh batch (k = 3, its hierarchy allowed 48 levels): 0.015 / 512
# takes 30 iterations (2,455 solves/s), 0.003 / 4096 17 (3,009 solves/s);
# with the FP32 coarse inverse 0.003 / 4096 / 5120 / 6144 / 8192 / 12288
# rows gave 3,130 / 3,154 / 3,350 / 3,313 / 3,188 solves/s (17, 17, 16, 16,
# 16 iterations), 0.0025 / 0.0035 / 0.004 at 6144 3,172 / 3,245 / 3,236.
HSC_SHAPE = {"default": (0.015, 512, 24), "capsules100k": (0.002, 4096, 24),
             "icosphere9": (0.0015, 12288, 48), "meshes64": (0.003, 6144, 24)}
COPY_KERNEL = "__amd_rocclr_copyBuffer"


def traffic_from(fetch_csv, write_csv, copy_mb, out=TRAFFIC_FILE):
    """HBM bytes per dispatch from two `rocprofv3 --pmc` runs (FETCH_SIZE,
    then WRITE_SIZE, KiB per dispatch), each scaled so that the plain
    `copy_mb` MiB device copy reads and writes exactly copy_mb MiB."""
    import csv

    def load(path, counter):
        per = {}
        for r in csv.DictReader(open(path)):
            if r["Counter_Name"] != counter:
                continue
            per.setdefault(r["Kernel_Name"], []).append(float(r["Counter_Value"]))
        return per

    f, w = load(fetch_csv, "FETCH_SIZE"), load(write_csv, "WRITE_SIZE")
    copy_kib = copy_mb * 1024.0
    cf = copy_kib / max(max(v) for k_, v in f.items() if k_.startswith(COPY_KERNEL))
    cw = copy_kib / max(max(v) for k_, v in w.items() if k_.startswith(COPY_KERNEL))
    res = {"workload": WORKLOAD, "copy_mib": copy_mb, "fetch_factor": cf, "write_factor": cw,
           "kernels": {}}
    for tag, prefixes in TRAFFIC_KERNELS.items():
        prefixes = (prefixes,) if isinstance(prefixes, str) else prefixes
        fk = wk = 0.0
        nd = None
        for prefix in prefixes:
            fv = [x for k_, v in f.items() if k_.startswith(prefix + "(") for x in v]
            wv = [x for k_, v in w.items() if k_.startswith(prefix + "(") for x in v]
            if not fv or not wv:
                break
            fk, wk = fk + sum(fv) / len(fv), wk + sum(wv) / len(wv)
            nd = len(fv) if nd is None else nd
        else:
            res["kernels"][tag] = {"kernel": " + ".join(prefixes), "dispatches": nd,
                                   "fetch_kib_raw": fk, "write_kib_raw": wk,
                                   "fetch_bytes": fk * 1024 * cf, "write_bytes": wk * 1024 * cw,
                                   "bytes": fk * 1024 * cf + wk * 1024 * cw}
    with open(out, "w") as fh:
        json.dump(res, fh, indent=1)
    print(json.dumps(res, indent=1))


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--poses", type=int, default=256)
    ap.add_argument("--tol", type=float, default=1e-9)
    ap.add_argument("--precond", default="hsc", choices=["hsc", "ic0"],
                    help="hierarchical Schur (default) or IC(0)")
    ap.add_argument("--theta", type=float, default=None,
                    help="HSC drop threshold (default: the workload's, see HSC_SHAPE)")
    ap.add_argument("--coarse-max", type=int, default=None,
                    help="HSC: stop eliminating at this many rows (dense coarse matrix)")
    ap.add_argument("--max-levels", type=int, default=None, help="HSC: most elimination levels")
    ap.add_argument("--inner-tol", type=float, default=None,
                    help="mixed precision: residual replacement once a column's iterated "
                         "||r|| has fallen by this factor (default: the autotuner's choice)")
    ap.add_argument("--ordering", default="block", choices=["block", "color", "natural"],
                    help="IC(0) ordering (block multicolour, point colour, natural levels)")
    ap.add_argument("--block-size", type=int, default=16)
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp64"],
                    help="inner CG arithmetic (FP32 with FP64 updates, or FP64)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-poses", type=int, default=8, help="poses in the CPU baseline sample")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="also run a plain device copy of this many MiB (counter calibration)")
    ap.add_argument("--plain-poses", type=int, default=8,
                    help="poses solved by the plain FP64 one-solve-per-launch path")
    ap.add_argument("--workload", default="capsules100k",
                    choices=["capsules100k", "meshes64", "icosphere9", "ddm", "setup"],
                    help="capsules100k: 256 poses x 3 RHS on the 100k capsule union (the "
                         "default); meshes64: 64 uneven meshes x 3 RHS, batched vs one solve "
                         "per mesh; icosphere9: 16 poses x 3 RHS forward + adjoint on the "
                         "2.6M-vertex icosphere with 64 loops; ddm: Direct Delta Mush bind "
                         "(harvest of every handle, smoothing, top-8) and the FP32 skinning "
                         "matvec over 256 poses on the capsule union; setup: the "
                         "set-up stages (seam cut, system build, hierarchy, RHS table) "
                         "and the re-assembly kernels on the capsule union and the "
                         "icosphere")
    ap.add_argument("--traffic-from", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="no GPU: turn two rocprofv3 --pmc counter files (FETCH_SIZE, "
                         "WRITE_SIZE) of a --copy-mb run into the traffic file")
    args = ap.parse_args()
    shape = HSC_SHAPE.get(args.workload, HSC_SHAPE["default"])
    for f, v in zip(("theta", "coarse_max", "max_levels"), shape):
        if getattr(args, f) is None:
            setattr(args, f, v)
    if args.traffic_from:
        traffic_from(*args.traffic_from, copy_mb=args.copy_mb or 1024)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "the solver runs on one GPU; --gpus must be 1 "
                          "(sharding across GPUs is not implemented)"}))
        sys.exit(2)

    import numpy as np
    import torch

    import tool_godot_curvenet_amd as cn
    from tool_godot_curvenet_amd import _lib, meshgen

    assert torch.cuda.is_available(), "no GPU"
    lib = _lib.load()
    dev = torch.device("cuda:0")

    if args.workload == "meshes64":
        meshes64_bench(args)
        return
    if args.workload == "icosphere9":
        icosphere9_bench(args)
        return
    if args.workload == "ddm":
        ddm_bench(args)
        return
    if args.workload == "setup":
        setup_bench(args)
        return

    if args.copy_mb > 0:
        # a plain copy of known size: its FETCH_SIZE / WRITE_SIZE counters
        # calibrate the HBM traffic counters of the solver kernels
        src = torch.ones(args.copy_mb * 262144, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    # ---------------------------------------------------------- setup ------
    t0 = time.time()
    mesh = meshgen.config_mesh("capsules100k")
    lap = cn.CutLaplacian.from_mesh(mesh)
    if args.precond == "hsc":
        pre = cn.HSC(lap, theta=args.theta, coarse_max=args.coarse_max,
                     max_levels=args.max_levels)
        pre_desc = (f"HSC theta {args.theta} ({pre.info.nlevels} levels, coarse "
                    f"{pre.info.n_coarse})")
    else:
        pre = cn.IC0(lap, args.ordering, block_size=args.block_size)
        pre_desc = (f"IC(0) {args.ordering} order ({pre.info.nlevels} colours, "
                    f"{pre.info.nblocks} blocks)")
    n, ncon = lap.n_free, lap.n_con
    nnz = lap.info.nnz_ff
    P = args.poses
    k = 3 * P
    rng = np.random.default_rng(meshgen.BASE_SEED + 30)
    Xc = torch.as_tensor(rng.uniform(-0.2, 0.2, size=(ncon, k)), device=dev)
    B = torch.empty((n, k), dtype=torch.float64, device=dev)
    X = torch.zeros((n, k), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    setup_s = time.time() - t0

    opts = _lib.PcgOpts()
    lib.cn_pcg_default_opts(opts)
    opts.precond = _lib.CN_PRECOND_IC0 if args.precond == "ic0" else _lib.CN_PRECOND_JACOBI
    opts.precision = _lib.CN_PREC_MIXED if args.precision == "mixed" else _lib.CN_PREC_FP64
    opts.tol_rel = args.tol
    stream = torch.cuda.current_stream().cuda_stream
    import ctypes as C
    sp = C.c_void_p(stream)

    def step(info):
        st = lib.cn_harmonic_rhs(lap.handle, k, C.c_void_p(Xc.data_ptr()),
                                 C.c_void_p(B.data_ptr()), sp)
        _lib.check(st, "harmonic_rhs")
        solve = lib.cn_pcg_solve_hsc if args.precond == "hsc" else lib.cn_pcg_solve
        st = solve(lap.handle, pre.handle, C.byref(opts), k,
                              C.c_void_p(B.data_ptr()), C.c_void_p(X.data_ptr()),
                              C.byref(info), sp)
        _lib.check(st, "pcg_solve")

    info = _lib.SolveInfo()
    # launch shapes: the autotuner times full solves of this workload's
    # right-hand sides for each candidate (setup, outside the timed region)
    _lib.check(lib.cn_harmonic_rhs(lap.handle, k, C.c_void_p(Xc.data_ptr()),
                                   C.c_void_p(B.data_ptr()), sp), "harmonic_rhs")
    tune = tune_for(lap, B, args, pre)
    opts.inner_tol = args.inner_tol if args.inner_tol is not None else tune["inner_tol"]
    for _ in range(args.warmup):
        step(info)
    torch.cuda.synchronize()

    PROF_NAMES = ("spmm", "spmm_lds", "spmm_dir", "spmm_upd", "fwd", "bwd", "resid", "dot",
                  "hsc_down", "hsc_up", "hsc_coarse", "hsc_tail")

    def run_steps(nsteps, only=None):
        """nsteps solves with HIP events around the kernels' launches (only:
        around the scopes of that name alone); returns the per-scope counts and
        milliseconds, the inner iterations, replacements, worst relative
        residual and the wall time."""
        lib.cn_profile_only(only.encode() if only else None)
        lib.cn_profile_enable(1)
        prof = {name: [0, 0.0] for name in PROF_NAMES}
        it = rf = 0
        rel = 0.0
        torch.cuda.synchronize()
        t = time.perf_counter()
        names = [only] if only else list(prof)  # the scopes that record events
        for _ in range(nsteps):
            step(info)
            for name in names:
                ms = C.c_double(0.0)
                cnt = lib.cn_profile_query(name.encode(), C.byref(ms))
                prof[name][0] += cnt
                prof[name][1] += ms.value
            it += info.iters
            rf += info.refinements
            rel = max(rel, info.relres_max)
        torch.cuda.synchronize()
        el = time.perf_counter() - t
        lib.cn_profile_enable(0)
        lib.cn_profile_only(None)
        return prof, it, rf, rel, el

    # Every kernel timed over a few untimed steps gives the breakdown and the
    # dominant kernel.  The timed steps then record events only around the
    # dominant kernel's launches: each event pair idles the GPU a few
    # microseconds, about 0.6 ms a step with every scope recorded.
    nbreak = max(1, min(args.steps, 3))
    prof_b, iters_b, refines_b, _, _ = run_steps(nbreak)
    fused = prof_b["spmm_dir"][0] > 0

    def kernel_table(prof, iters, refines, nsteps):
        """Per-kernel launches, time and algorithmic bytes (or flops) of
        nsteps steps timed by run_steps; returns the kernel table, the coarse
        product's entry and the profiling scope of each kernel."""
        # ---- roofline of the kernels of the inner loop (HIP events around
        # every launch of the steps timed).  Algorithmic bytes: the SpMM reads
        # A (FP32 value + int32 column per nonzero, row pointers) and P and
        # writes Q; a fused forward sweep reads R, Q and writes R, Y (4 n x k
        # FP32 vectors), the plain one after each residual replacement reads R
        # and writes Y (2); a backward sweep reads Y, P, D and writes Z, P, D
        # (6).  Sweeps launch once per colour level; gathers of rows from other
        # colours and the index tables are not counted.  Launches that find
        # every column converged return at once and are given no bytes.
        nk4 = 4 * n * k
        spmm_bytes = 8 * nnz + 4 * (n + 1) + 2 * nk4
        # The SpMM runs as the fused pair of LDS-staged products (KL_DIR: the
        # direction update p = beta p + gamma z with q = A p and p.q -- reads
        # P and Z, writes P; KL_UPD: q = A p again with r -= alpha q, d += alpha
        # p and r.r -- reads P, R, D, writes R, D), or as one product with Q
        # stored (LDS-staged or row-table) followed by kw_rupd and kw_pd_update.
        if fused:
            dir_bytes = 8 * nnz + 4 * (n + 1) + 3 * nk4
            upd_bytes = 8 * nnz + 4 * (n + 1) + 5 * nk4
            kern = {
                "spmm_dir": ("kl_spmm<float,6,KL_DIR>", prof["spmm_dir"][0],
                             prof["spmm_dir"][1], iters * dir_bytes),
                "spmm_upd": ("kl_spmm<float,6,KL_UPD>", prof["spmm_upd"][0],
                             prof["spmm_upd"][1], iters * upd_bytes),
            }
        else:
            sp_tag = "spmm_lds" if prof["spmm_lds"][0] > 0 else "spmm"
            sp_name = "kl_spmm<float,6,KL_PLAIN>" if sp_tag == "spmm_lds" else \
                "kw_spmm_dot<float,4,64>"
            kern = {
                "spmm": (sp_name, prof[sp_tag][0], prof[sp_tag][1], iters * spmm_bytes),
            }
        coarse = None
        mfma = {}
        if args.precond == "ic0":
            nlev = pre.info.nlevels
            kern["fwd"] = ("kb_fwd<float,4,16>", prof["fwd"][0] * nlev, prof["fwd"][1],
                           (4 * iters + 2 * refines) * nk4)
            kern["bwd"] = ("kb_bwd<float,4,16>", prof["bwd"][0] * nlev, prof["bwd"][1],
                           6 * (iters + refines) * nk4)
        else:
            # HSC apply: the down sweep as down_sweep_bytes counts it (each row
            # read and written once, distinct gathered rows once per launch); up,
            # per level l (n_F eliminated, n_C kept rows, k FP32 columns), reads
            # and writes the eliminated rows and reads, once, the kept ones:
            # 4k(2 n_F + n_C) plus P, the row map and 1/a_ff.
            # kw_pd_update (d += alpha p, p = z + beta p) streams 5 n x k vectors.
            # levels [l0, L) and the coarse solve run in the fused tail launch
            l0 = lib.cn_hsc_tail_level(pre.handle, k, 1 if args.precision == "mixed" else 0)
            levels = pre.levels()
            down_b = up_b = tail_b = 0
            for li, lv in enumerate(levels):
                nF, nC = len(lv["F"]), len(lv["C"])
                db = 4 * k * (2 * nC + nF) + 8 * len(lv["R"][1]) + 4 * (nC + 1) + 4 * nC
                ub = 4 * k * (2 * nF + nC) + 8 * len(lv["P"][1]) + 4 * (nF + 1) + 8 * nF
                if 0 <= l0 <= li:
                    tail_b += db + ub
                else:
                    up_b += ub
            down_b = down_sweep_bytes(levels, n, k, l0 if l0 >= 0 else len(levels),
                                      lap.get_tune()["chain_levels"] or 3)
            # Bytes are counted for the applies known to do work: one per active
            # iteration and the first of each solve (a restart adds an uncounted
            # one; launches during a pause return at once).  Time covers every
            # launch, so the achieved rates are lower bounds.
            nlev = max(pre.info.nlevels if l0 < 0 else l0, 1)
            nwork = iters + nsteps
            # a sweep's scope encloses its nlev launches of an apply
            kern["hsc_down"] = ("kw_hsc_chain_rec<float,4,64>", prof["hsc_down"][0] * nlev,
                                prof["hsc_down"][1], nwork * down_b)
            kern["hsc_up"] = ("kw_hsc_up<float,4,64>", prof["hsc_up"][0] * nlev, prof["hsc_up"][1],
                              nwork * up_b)
            if l0 >= 0:
                # the tail: its levels' down and up bytes, the coarse rows read and
                # written, and the FP64 coarse inverse once per apply
                ncr = pre.info.n_coarse
                tail_b += 8 * k * ncr + 8 * ncr * ncr
                kern["hsc_tail"] = ("k_hsc_tail<float,4>", prof["hsc_tail"][0], prof["hsc_tail"][1],
                                    nwork * tail_b)
            # kw_pd_update: every iteration's direction update, or with the fused
            # products only the deferred update before a round that re-applies
            # the preconditioner
            nbwd = prof["bwd"][0]
            kern["pd_update"] = ("kw_pd_update<float,4,64>", nbwd, prof["bwd"][1],
                                 (nbwd if fused else nwork) * 5 * nk4)
            # r.z where the apply does not form it from its passes: reads R and Z
            kern["dot"] = ("kw_dot<float,4,64>", prof["dot"][0], prof["dot"][1],
                           prof["dot"][0] * 2 * nk4)
            nc = pre.info.n_coarse
            cms = prof["hsc_coarse"][1]
            ncoarse = prof["hsc_coarse"][0]
            # the coarse product on FP32 (coarse_prec 2), split-bfloat16
            # (coarse_prec 3, hierarchies of at least 3,072 coarse rows) or
            # FP64 matrix cores
            cp = lap.get_tune()["coarse_prec"] if args.precision == "mixed" else 1
            cbf = cp == 3 and nc >= 3072 and k > 4
            c32 = cp >= 2 and not cbf
            cprec, cpeak = (("split bfloat16, 3 products", BF16_MFMA_PEAK_TFLOPS) if cbf else
                            ("FP32", FP32_MFMA_PEAK_TFLOPS) if c32 else
                            ("FP64", FP64_MFMA_PEAK_TFLOPS))
            ctag = "hsc_coarse_bf3" if cbf else "hsc_coarse" if c32 else "hsc_coarse_f64"
            if l0 < 0 and ncoarse > 0:
                # the coarse product: 2 n_c^2 k flops per apply that does work
                # (the split product runs three times as many on the MFMAs)
                big = nc >= 3072 and k >= 512
                kn = (f"k_coarse_mm_bf3<float,{'128,128,4,2' if big else '64,64,2,2'}>" if cbf else
                      f"k_coarse_mm<float,{'float' if c32 else 'double'},128,128,4,2>")
                mfma[ctag] = (f"{kn} + k_coarse_sum ({cprec} MFMA)", ncoarse, cms,
                              nwork * 2.0 * nc * nc * k, cpeak)
            tflops = 2.0 * nc * nc * k * nwork / (cms * 1e-3) / 1e12 if cms > 0 else 0.0
            if l0 >= 0:
                coarse = {"kernel": "coarse solve inside k_hsc_tail (FP64 on the vector units, "
                                    "one workgroup per column slice)", "n_coarse": nc,
                          "tail_from_level": int(l0), "levels": int(pre.info.nlevels)}
            else:
                coarse = {"kernel": f"k_coarse_mm (coarse inverse x coarse rows, {cprec} MFMA)",
                          "n_coarse": nc, "launches": ncoarse, "avg_ms": cms / max(ncoarse, 1),
                          "achieved_tflops": tflops, "peak_tflops": cpeak,
                          "frac": tflops / cpeak}
        kernels = {}
        for tag, (kname, launches, ms, nbytes) in kern.items():
            gbs = nbytes / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
            kernels[tag] = {"kernel": kname, "bound": "hbm", "launches": launches, "ms": ms,
                            "avg_ms": ms / max(launches, 1),
                            "bytes_per_launch": nbytes / max(launches, 1),
                            "achieved_gbs": gbs, "frac": gbs / HBM_PEAK_GBS}
        for tag, (kname, launches, ms, flops, peak) in mfma.items():
            tf = flops / (ms * 1e-3) / 1e12 if ms > 0 else 0.0
            kernels[tag] = {"kernel": kname, "bound": "mfma", "launches": launches, "ms": ms,
                            "avg_ms": ms / max(launches, 1),
                            "flops_per_launch": flops / max(launches, 1),
                            "achieved_tflops": tf, "peak": peak, "frac": tf / peak}
        # kernels this workload does not launch (the fused iteration has no
        # kw_pd_update; the apply forms r.z itself) are left out
        kernels = {t: v for t, v in kernels.items() if v["launches"] > 0}
        scope = {t: t for t in kernels}
        scope["pd_update"] = "bwd"
        if "spmm" in kernels:
            scope["spmm"] = "spmm_lds" if prof["spmm_lds"][0] > 0 else "spmm"
        for t in ("hsc_coarse", "hsc_coarse_f64", "hsc_coarse_bf3"):
            scope[t] = "hsc_coarse"
        return kernels, coarse, scope

    kernels, coarse, scope = kernel_table(prof_b, iters_b, refines_b, nbreak)
    dom = max(kernels, key=lambda t: kernels[t]["ms"])
    prof, iters, refines, relmax, elapsed = run_steps(args.steps, only=scope[dom])
    kt, coarse_t, _ = kernel_table(prof, iters, refines, args.steps)
    kernels[dom] = {**kt[dom], "timed_region": True}
    if scope[dom] == "hsc_coarse":
        coarse = coarse_t

    # independent FP64 residual check of the last step (not timed)
    A = lap.A_ff()
    import scipy.sparse as sps
    Acsr = sps.csr_matrix((A.values, A.col_idx, A.row_ptr), shape=(n, n))
    Xh, Bh = X.cpu().numpy(), B.cpu().numpy()
    res = np.linalg.norm(Bh - Acsr @ Xh, axis=0) / np.linalg.norm(Bh, axis=0)
    check_relres = float(res.max())

    solves = P * args.steps
    value = solves / elapsed

    # --------------------------------------- plain FP64 one-solve-per-launch
    # The unfused FP64 path (one launch per SpMV / dot / axpy / preconditioner
    # step, alpha and beta on the host), one pose (3 RHS) per solve, with the
    # same HSC hierarchy as the headline (the same-preconditioner baseline),
    # with block IC(0), and with Jacobi.
    plain = None
    if args.plain_poses > 0:
        Bp = B[:, :3].contiguous()
        Xp = torch.zeros_like(Bp)
        plain = {}
        pres = [("hsc", pre if args.precond == "hsc" else
                 cn.HSC(lap, theta=args.theta, coarse_max=args.coarse_max,
                        max_levels=args.max_levels)),
                ("ic0", pre if args.precond == "ic0" else cn.IC0(lap, "block", block_size=16)),
                ("jacobi", None)]
        for pname, pobj in pres:
            popts = _lib.PcgOpts()
            lib.cn_pcg_default_opts(popts)
            popts.precond = {"ic0": _lib.CN_PRECOND_IC0}.get(pname, _lib.CN_PRECOND_JACOBI)
            popts.precision = _lib.CN_PREC_FP64_PLAIN
            popts.tol_rel = args.tol
            popts.max_iter = 200000
            pinfo = _lib.SolveInfo()
            fn = lib.cn_pcg_solve_hsc if pname == "hsc" else lib.cn_pcg_solve
            hdl = pobj.handle if pobj is not None else None
            npz = args.plain_poses if pname != "jacobi" else min(args.plain_poses, 2)
            fn(lap.handle, hdl, C.byref(popts), 3, C.c_void_p(Bp.data_ptr()),
               C.c_void_p(Xp.data_ptr()), C.byref(pinfo), sp)          # warm-up
            torch.cuda.synchronize()
            tp = time.perf_counter()
            pit = 0
            for _ in range(npz):
                fn(lap.handle, hdl, C.byref(popts), 3, C.c_void_p(Bp.data_ptr()),
                   C.c_void_p(Xp.data_ptr()), C.byref(pinfo), sp)
                pit += pinfo.iters
            torch.cuda.synchronize()
            tpe = time.perf_counter() - tp
            plain[pname] = {"solves_per_s": npz / tpe, "iters_per_solve": pit / npz,
                            "relres_max": pinfo.relres_max, "status": pinfo.status,
                            "poses": npz}

    # ---------------------------------------------------------- CPU leg ----
    cpu = None
    if not args.no_cpu:
        import oracle as O
        lp, lc, lv = O.ic0_factor(A.row_ptr, A.col_idx, A.values)
        fac = (lp, lc, lv)
        # (i) one thread, the reference's sequential semantics
        npose = max(1, min(args.cpu_poses, P))
        tc = time.perf_counter()
        _, cit = O.pcg_many(A.row_ptr, A.col_idx, A.values, Bh[:, :3 * npose], args.tol,
                            factor=fac, threads=1)
        tce = time.perf_counter() - tc
        # (ii) the columns spread over the host cores this job may use
        threads = max(1, min(16, os.cpu_count() or 1))
        mpose = max(1, min(P, npose * threads // 2))
        tm = time.perf_counter()
        _, mit = O.pcg_many(A.row_ptr, A.col_idx, A.values, Bh[:, :3 * mpose], args.tol,
                            factor=fac, threads=threads)
        tme = time.perf_counter() - tm
        cpu = {"value": mpose / tme, "unit": "solves/s", "cores": threads, "kind": "port",
               "sample": f"{mpose} poses x 3 RHS of the same workload, FP64 IC(0)-PCG natural "
                         f"order (factor outside the timing), columns over {threads} threads, "
                         f"{mit} iterations, {tme:.1f} s; host: {_cpu_model()}, "
                         f"{os.cpu_count()} cores visible",
               "single_thread": {"value": npose / tce, "unit": "solves/s", "cores": 1,
                                 "sample": f"{npose} poses x 3 RHS, {cit} iterations, "
                                           f"{tce:.1f} s"}}

    # ---- k = 1 FP32 SpMV of the same matrix (cn_spmv_f32), HIP events on
    # the launch stream.  Algorithmic bytes: 8 nnz (FP32 value + int32
    # column) + 4 (n + 1) row pointers + 4 n (x) + 4 n (y).  The 5.9 MB
    # working set stays in the 256 MB Infinity Cache, so this is an on-chip
    # rate, not an HBM one (the level-9 icosphere workload gives the HBM one).
    spmv = spmv_k1(lib, lap.A_ff(), sp, reps=200)

    # ---- C1: the reference's own CPU solver (DenseLinAlg.solveMulti: dense
    # Gaussian elimination with partial pivoting, re-run per column) on the
    # 64 x 64 cut cylinder, one column timed, beside the GPU solve of the
    # same system (IC(0)-PCG to 1e-10) and the difference of the two.
    c1 = None if args.no_cpu else c1_dense_vs_gpu(lib, sp)

    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as fh:
            tk = json.load(fh).get("kernels", {}).get(dom)
        same = (args.precond == "hsc" and
                (args.theta, args.coarse_max, args.max_levels) == HSC_SHAPE[WORKLOAD]) or \
            (args.precond == "ic0" and args.ordering == "block" and args.block_size == 16)
        if tk and k == 768 and same:
            traffic = tk["bytes"]

    kd = kernels[dom]
    if kd["bound"] == "mfma":
        roofline = {"bound": "mfma", "achieved": kd["achieved_tflops"],
                    "peak": kd["peak"], "unit": "TFLOP/s", "frac": kd["frac"],
                    "traffic": traffic, "kernel": kd["kernel"], "avg_ms": kd["avg_ms"],
                    "launches": kd["launches"], "flops_per_launch": kd["flops_per_launch"]}
    else:
        roofline = {"bound": "hbm", "achieved": kd["achieved_gbs"], "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": kd["frac"], "traffic": traffic,
                    "kernel": kd["kernel"], "avg_ms": kd["avg_ms"], "launches": kd["launches"],
                    "bytes_per_launch": kd["bytes_per_launch"]}
    line = {
        "metric": "solves/s at 1e-9 rel. residual (100k-vert cut Laplacian, 3 RHS)",
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("f32 inner / f64 refinement" + (
            " (coarse product: bf16 MFMA on hi + lo splits, 2^-16)"
            if args.precond == "hsc" and lap.get_tune()["coarse_prec"] == 3 and
            pre.info.n_coarse >= 3072 and k > 4 else "")
            if args.precision == "mixed" else "f64"),
        "data": "synthetic (seeded capsule union, seeded constraint displacements)",
        "config": {"workload": "capsule union 12 capsules / 16 ring cuts, "
                               f"{mesh.nv} verts, n_free={n}, nnz={nnz}, "
                               f"{P} poses x 3 RHS (k={k}), {pre_desc}, "
                               f"tol {args.tol}"},
        "roofline": roofline,
        "kernels": kernels,
        "tune": tune,
        "spmv_k1": spmv,
        "c1_dense_ge": c1,
        "coarse_mfma": coarse,
        "cpu_baseline": cpu,
        "plain_fp64_gpu": plain,
        # every scope timed over the untimed breakdown steps
        "kernel_ms_per_step": {k_: v[1] / nbreak for k_, v in prof_b.items()},
        "breakdown_steps": nbreak,
        "solve": {"inner_iters_per_step": iters / args.steps,
                  "refinements_per_step": refines / args.steps,
                  "relres_max": relmax, "relres_check_scipy": check_relres,
                  "setup_s": setup_s},
    }
    if plain:
        # the same preconditioner as the headline on both sides
        line["speedup_vs_plain_fp64"] = value / plain[args.precond]["solves_per_s"]
        line["speedup_vs_plain_fp64_jacobi"] = value / plain["jacobi"]["solves_per_s"]
    print(json.dumps(line))


def down_sweep_bytes(levels, n0, k, t, G=3):
    """Algorithmic bytes of one HSC down sweep up to level t in groups of G
    levels (kw_hsc_chain_rec): in group [a, e), launch m (a < m < e) takes the
    rows eliminated at level m, the last launch the rows of level e, each
    from its level-a value through the R entries of levels a .. m-1; each
    row is read and written once (4k bytes each way), every distinct row
    its chain gathers is read once per launch, each chain entry costs its
    column and FP32 weight (8 B), plus the row list, segment pointer and
    three segment bounds (20 B a row)."""
    import numpy as np
    L = len(levels)
    if t <= 0 or L == 0:
        return 0
    g = np.arange(n0, dtype=np.int64)           # level-l row -> level-0 row
    cls = np.full(n0, L, np.int64)              # level each row is eliminated at
    ent_l, ent_v, ent_u = [], [], []
    for l, lv in enumerate(levels):
        F, C = lv["F"].astype(np.int64), lv["C"].astype(np.int64)
        cls[g[F]] = l
        rp, rc, _ = lv["R"]
        rows = np.repeat(np.arange(len(C)), np.diff(rp))
        ent_l.append(np.full(len(rc), l, np.int64))
        ent_v.append(g[C][rows])
        ent_u.append(g[rc.astype(np.int64)])
        g = g[C]
    el, ev, eu = (np.concatenate(a) for a in (ent_l, ent_v, ent_u))
    total, a = 0, 0
    while a < t:
        e = min(a + G, t)
        for m in range(a + 1, e + 1):
            rows = (cls == m) if m < e else (cls >= e)
            nrows = int(rows.sum())
            sel = (el >= a) & (el < m) & rows[ev]
            distinct = int(np.unique(eu[sel]).size)
            total += 4 * k * (2 * nrows + distinct) + 8 * int(sel.sum()) + 20 * nrows
        a = e
    return total


def tune_for(system, B, args, pre, precond=None):
    """Autotune the solver's launch shapes on this workload (no cache: the
    box is timed as it is) and report the choice beside the built-in one."""
    from tool_godot_curvenet_amd import autotune
    precond = precond or args.precond
    kw = {"hsc": pre} if precond == "hsc" else {"ic0": pre}
    r = autotune.autotune(system, B, precond=precond, precision=args.precision, tol=args.tol,
                          use_cache=False, **kw)
    built_in = next((c["ms"] for c in r["candidates"]
                     if c["tune"] == {**autotune.DEFAULT, "spmm_form": 2}
                     and c["inner_tol"] == autotune.INNER_TOLS[0]), None)
    tried = [{"set": {f: v for f, v in c["tune"].items() if v != autotune.DEFAULT[f]},
              "inner_tol": c["inner_tol"], "ms": round(c["ms"], 3)} for c in r["candidates"]]
    return {"chosen": r["tune"], "inner_tol": r["inner_tol"], "solve_ms": r["ms"],
            "built_in_solve_ms": built_in, "candidates": len(r["candidates"]),
            "tried": tried,
            "how": "coordinate search over SpMM form / stripes, HSC level-pass grid cap, "
                   "fused-tail size, down-sweep chain length and grid cap, coarse product "
                   "precision, "
                   "residual-replacement drop; full solves of this workload, best of 2"}


def spmv_k1(lib, A, sp, reps=200):
    """Time cn_spmv_f32 on CSR A (FP32 values) over `reps` launches with HIP
    events on the launch stream; achieved GB/s of algorithmic bytes."""
    import ctypes as C

    import numpy as np
    import torch
    dev = torch.device("cuda:0")
    n, nnz = A.rows, A.nnz
    ptr = torch.as_tensor(A.row_ptr.astype(np.int32), device=dev)
    col = torch.as_tensor(A.col_idx.astype(np.int32), device=dev)
    val = torch.as_tensor(A.values.astype(np.float32), device=dev)
    x = torch.as_tensor(np.random.default_rng(1).normal(size=n).astype(np.float32), device=dev)
    y = torch.empty_like(x)
    args_ = (n, C.c_void_p(ptr.data_ptr()), C.c_void_p(col.data_ptr()),
             C.c_void_p(val.data_ptr()), C.c_void_p(x.data_ptr()), C.c_void_p(y.data_ptr()), sp)
    for _ in range(10):
        lib.cn_spmv_f32(*args_)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        lib.cn_spmv_f32(*args_)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    import scipy.sparse as sps
    ref = sps.csr_matrix((A.values.astype(np.float32).astype(np.float64), A.col_idx,
                          A.row_ptr), shape=(n, n)) @ x.cpu().numpy().astype(np.float64)
    err = float(np.abs(y.cpu().numpy() - ref).max() / max(np.abs(ref).max(), 1e-300))
    alg = 8 * nnz + 4 * (n + 1) + 8 * n
    gbs = alg / (ms * 1e-3) / 1e9
    return {"kernel": "kt_spmm<float,1> (cn_spmv_f32)", "n": n, "nnz": nnz, "avg_ms": ms,
            "bytes_per_launch": alg, "achieved_gbs": gbs, "frac": gbs / HBM_PEAK_GBS,
            "max_rel_err_vs_fp64": err}


def c1_dense_vs_gpu(lib, sp):
    import numpy as np
    import torch

    import oracle as O
    import tool_godot_curvenet_amd as cn
    from tool_godot_curvenet_amd import meshgen
    lap = cn.CutLaplacian.from_mesh(meshgen.config_mesh("cylinder64"))
    A = lap.A_ff()
    B = lap.harmonic_rhs(meshgen.seeded_boundary(lap.n_con, 3))
    ic0 = cn.IC0(lap)
    lap.solve(B, precond="ic0", precision="mixed", tol=1e-10, ic0=ic0)      # warm-up
    torch.cuda.synchronize()
    reps = 20
    t = time.perf_counter()
    for _ in range(reps):
        x, info = lap.solve(B, precond="ic0", precision="mixed", tol=1e-10, ic0=ic0)
    torch.cuda.synchronize()
    gpu_s = (time.perf_counter() - t) / reps
    D = A.dense()
    tc = time.perf_counter()
    xd = O.dense_solve(D, B[:, 0])
    col_s = time.perf_counter() - tc
    X = x.cpu().numpy()
    diff = float(np.abs(X[:, 0] - xd).max() / np.abs(xd).max())
    return {"n": A.rows, "cpu_kind": "port (oracle dense GE, one thread)",
            "cpu_seconds_per_column": col_s, "cpu_solves_per_s": 1.0 / (3 * col_s),
            "gpu_solves_per_s": 1.0 / gpu_s, "gpu_iters": info.iters,
            "gpu_vs_dense_max_rel_diff": diff, "cpu": _cpu_model()}


def meshes64_bench(args):
    """64 capsule unions of log-uniform size (2k-200k vertices), 3 RHS
    each, solved to 1e-9: all of them as one batch (one launch per solver
    step covers every mesh, HSC on the block-diagonal system) against one
    solve per mesh (each with its own HSC) on the same inputs."""
    import numpy as np
    import scipy.sparse as sps
    import torch

    import tool_godot_curvenet_amd as cn
    from tool_godot_curvenet_amd import _lib, meshgen
    import ctypes as C

    assert torch.cuda.is_available(), "no GPU"
    lib = _lib.load()
    dev = torch.device("cuda:0")
    t0 = time.time()
    meshes = meshgen.meshes64()
    print(f"meshes64: {len(meshes)} meshes, {sum(m.nv for m in meshes)} vertices generated "
          f"in {time.time() - t0:.1f} s", file=sys.stderr, flush=True)
    laps = [cn.CutLaplacian.from_mesh(m) for m in meshes]
    batch = cn.BatchSystem(laps)
    hb = cn.HSC(batch, theta=args.theta, coarse_max=args.coarse_max,
                max_levels=max(args.max_levels, 48))
    hs = [cn.HSC(l, theta=args.theta, coarse_max=args.coarse_max,
                 max_levels=args.max_levels) for l in laps]
    print(f"meshes64: systems and hierarchies built at {time.time() - t0:.1f} s",
          file=sys.stderr, flush=True)
    k = 3
    N = batch.n_free
    rng = np.random.default_rng(meshgen.BASE_SEED + 40)
    Xc = [torch.as_tensor(rng.uniform(-0.2, 0.2, size=(l.n_con, k)), device=dev) for l in laps]
    Bb = torch.empty((N, k), dtype=torch.float64, device=dev)
    Xb = torch.zeros_like(Bb)
    Bs = [torch.empty((l.n_free, k), dtype=torch.float64, device=dev) for l in laps]
    Xs = [torch.zeros_like(b) for b in Bs]
    seg = batch.seg_ptr
    sp = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    opts = _lib.PcgOpts()
    lib.cn_pcg_default_opts(opts)
    opts.precision = _lib.CN_PREC_MIXED if args.precision == "mixed" else _lib.CN_PREC_FP64
    opts.tol_rel = args.tol
    info = _lib.SolveInfo()

    Xcs = torch.cat(Xc, dim=0).contiguous()      # the members' constrained values, stacked
    opts_direct = _lib.PcgOpts()
    C.memmove(C.byref(opts_direct), C.byref(opts), C.sizeof(opts))
    opts_direct.graph = 0

    def step_batch():
        # every member's right-hand side in one launch, then the batched
        # solve with its inner iterations replayed from a captured graph
        _lib.check(lib.cn_harmonic_rhs(batch.handle, k, C.c_void_p(Xcs.data_ptr()),
                                       C.c_void_p(Bb.data_ptr()), sp), "harmonic_rhs")
        _lib.check(lib.cn_pcg_solve_hsc(batch.handle, hb.handle, C.byref(opts), k,
                                        C.c_void_p(Bb.data_ptr()), C.c_void_p(Xb.data_ptr()),
                                        C.byref(info), sp), "pcg_solve_hsc")
        return info.iters

    def step_batch_direct():
        # the same batch with one right-hand side launch per member and every
        # solver kernel launched directly (no graph)
        for i, l in enumerate(laps):
            _lib.check(lib.cn_harmonic_rhs(l.handle, k, C.c_void_p(Xc[i].data_ptr()),
                                           C.c_void_p(Bb[seg[i]:seg[i + 1]].data_ptr()), sp),
                       "harmonic_rhs")
        _lib.check(lib.cn_pcg_solve_hsc(batch.handle, hb.handle, C.byref(opts_direct), k,
                                        C.c_void_p(Bb.data_ptr()), C.c_void_p(Xb.data_ptr()),
                                        C.byref(info), sp), "pcg_solve_hsc")
        return info.iters

    def step_each():
        it = 0
        for i, l in enumerate(laps):
            _lib.check(lib.cn_harmonic_rhs(l.handle, k, C.c_void_p(Xc[i].data_ptr()),
                                           C.c_void_p(Bs[i].data_ptr()), sp), "harmonic_rhs")
            _lib.check(lib.cn_pcg_solve_hsc(l.handle, hs[i].handle, C.byref(opts), k,
                                            C.c_void_p(Bs[i].data_ptr()),
                                            C.c_void_p(Xs[i].data_ptr()), C.byref(info), sp),
                       "pcg_solve_hsc")
            it = max(it, info.iters)
        return it

    def timed(fn):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        its = [fn() for _ in range(args.steps)]
        torch.cuda.synchronize()
        return time.perf_counter() - t, its

    _lib.check(lib.cn_harmonic_rhs(batch.handle, k, C.c_void_p(Xcs.data_ptr()),
                                   C.c_void_p(Bb.data_ptr()), sp), "harmonic_rhs")
    tune = tune_for(batch, Bb, args, hb, precond="hsc")
    opts.inner_tol = opts_direct.inner_tol = \
        args.inner_tol if args.inner_tol is not None else tune["inner_tol"]
    ed, _ = timed(step_batch_direct)
    eb, itb = timed(step_batch)
    print(f"meshes64: batched {len(laps) * args.steps / eb:.1f} solves/s (direct launches "
          f"{len(laps) * args.steps / ed:.1f})", file=sys.stderr, flush=True)
    ee, ite = timed(step_each)
    # FP64 residual of every mesh, both paths, on the CPU (not timed)
    worst = [0.0, 0.0]
    Xbh, Bbh = Xb.cpu().numpy(), Bb.cpu().numpy()
    for i, l in enumerate(laps):
        A = l.A_ff()
        M = sps.csr_matrix((A.values, A.col_idx, A.row_ptr), shape=(l.n_free, l.n_free))
        for j, (x, b) in enumerate(((Xbh[seg[i]:seg[i + 1]], Bbh[seg[i]:seg[i + 1]]),
                                    (Xs[i].cpu().numpy(), Bs[i].cpu().numpy()))):
            r = np.linalg.norm(b - M @ x, axis=0) / np.linalg.norm(b, axis=0)
            worst[j] = max(worst[j], float(r.max()))
    value = len(laps) * args.steps / eb
    each = len(laps) * args.steps / ee
    sizes = [l.n_free for l in laps]
    line = {
        "metric": "solves/s at 1e-9 rel. residual (64 uneven meshes, 3 RHS each)",
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * eb / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 inner / f64 refinement" if args.precision == "mixed" else "f64",
        "data": "synthetic (64 seeded capsule unions, seeded constraint displacements)",
        "config": {"workload": f"64 capsule unions, n_free {min(sizes)}-{max(sizes)} "
                               f"(sum {N}), nnz {batch.info.nnz_ff}, 3 RHS each, HSC theta "
                               f"{args.theta} ({hb.info.nlevels} levels, coarse "
                               f"{hb.info.n_coarse}), tol {args.tol}"},
        "one_solve_per_mesh": {"solves_per_s": each, "ms_per_step": 1e3 * ee / args.steps,
                               "max_iters": max(ite)},
        "speedup_vs_one_solve_per_mesh": value / each,
        "tune": tune,
        "batched_direct_launches": {"solves_per_s": len(laps) * args.steps / ed,
                                    "ms_per_step": 1e3 * ed / args.steps,
                                    "what": "64 right-hand side launches, no graph replay"},
        "solve": {"inner_iters_per_step": sum(itb) / len(itb),
                  "relres_max_batched": worst[0], "relres_max_per_mesh": worst[1],
                  "setup_s": time.time() - t0 - eb - ee - ed},
    }
    print(json.dumps(line))


def icosphere9_bench(args):
    """The large-mesh workload: level-9 icosphere (2,621,442 vertices, 64
    tilted loops), 16 poses x 3 RHS; one step = harmonic RHS, forward solve
    to 1e-9 and the adjoint of a quadratic loss (lambda and dL/dx_c), both
    HSC-preconditioned, FP32 inner loop with FP64 refinement."""
    import numpy as np
    import scipy.sparse as sps
    import torch

    import tool_godot_curvenet_amd as cn
    from tool_godot_curvenet_amd import _lib, meshgen
    import ctypes as C

    assert torch.cuda.is_available(), "no GPU"
    lib = _lib.load()
    dev = torch.device("cuda:0")
    t0 = time.time()
    mesh = meshgen.config_mesh("icosphere9")
    print(f"icosphere9: {mesh.nv} vertices, {len(mesh.meta['loops'])} loops in "
          f"{time.time() - t0:.1f} s", file=sys.stderr, flush=True)
    lap = cn.CutLaplacian.from_mesh(mesh)
    hsc = cn.HSC(lap, theta=args.theta, coarse_max=args.coarse_max, max_levels=args.max_levels)
    print(f"icosphere9: system and hierarchy at {time.time() - t0:.1f} s", file=sys.stderr,
          flush=True)
    P = 16
    k = 3 * P
    n, ncon = lap.n_free, lap.n_con
    rng = np.random.default_rng(meshgen.BASE_SEED + 50)
    Xc = torch.as_tensor(rng.uniform(-0.2, 0.2, size=(ncon, k)), device=dev)
    target = torch.as_tensor(rng.normal(0, 0.05, size=(n, k)), device=dev)
    B = torch.empty((n, k), dtype=torch.float64, device=dev)
    X = torch.zeros_like(B)
    G = torch.empty_like(B)
    LAM = torch.empty_like(B)
    GC = torch.empty((ncon, k), dtype=torch.float64, device=dev)
    sp = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    opts = _lib.PcgOpts()
    lib.cn_pcg_default_opts(opts)
    opts.precision = _lib.CN_PREC_MIXED if args.precision == "mixed" else _lib.CN_PREC_FP64
    opts.tol_rel = args.tol
    fi, ai = _lib.SolveInfo(), _lib.SolveInfo()
    tf = [0.0]

    def step():
        _lib.check(lib.cn_harmonic_rhs(lap.handle, k, C.c_void_p(Xc.data_ptr()),
                                       C.c_void_p(B.data_ptr()), sp), "harmonic_rhs")
        _lib.check(lib.cn_pcg_solve_hsc(lap.handle, hsc.handle, C.byref(opts), k,
                                        C.c_void_p(B.data_ptr()), C.c_void_p(X.data_ptr()),
                                        C.byref(fi), sp), "pcg_solve_hsc")
        torch.sub(X, target, out=G)              # dL/dx for L = |x - target|^2 / 2
        _lib.check(lib.cn_adjoint_solve_hsc(lap.handle, hsc.handle, C.byref(opts_adj), k,
                                            C.c_void_p(G.data_ptr()), None, None,
                                            C.c_void_p(LAM.data_ptr()),
                                            C.c_void_p(GC.data_ptr()), None, C.byref(ai), sp),
                   "adjoint_solve_hsc")
        return fi.iters, ai.iters

    _lib.check(lib.cn_harmonic_rhs(lap.handle, k, C.c_void_p(Xc.data_ptr()),
                                   C.c_void_p(B.data_ptr()), sp), "harmonic_rhs")
    tune = tune_for(lap, B, args, hsc, precond="hsc")
    opts.inner_tol = args.inner_tol if args.inner_tol is not None else tune["inner_tol"]
    # the adjoint's right-hand side (x - target) is rougher than the forward
    # one: its residual-replacement drop is tuned on it (the launch shapes
    # stay the forward's)
    from tool_godot_curvenet_amd import autotune
    _lib.check(lib.cn_pcg_solve_hsc(lap.handle, hsc.handle, C.byref(opts), k,
                                    C.c_void_p(B.data_ptr()), C.c_void_p(X.data_ptr()),
                                    C.byref(fi), sp), "pcg_solve_hsc")
    torch.sub(X, target, out=G)
    adj = autotune.autotune(lap, G, precond="hsc", precision=args.precision, tol=args.tol,
                            use_cache=False, shapes=False, hsc=hsc)
    tune["adjoint_inner_tol"] = adj["inner_tol"]
    opts_adj = _lib.PcgOpts()
    C.memmove(C.byref(opts_adj), C.byref(opts), C.sizeof(opts))
    opts_adj.inner_tol = args.inner_tol if args.inner_tol is not None else adj["inner_tol"]
    free0, total = torch.cuda.mem_get_info()
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    its = [step() for _ in range(args.steps)]
    torch.cuda.synchronize()
    el = time.perf_counter() - t1
    # device memory in use (every allocation of the process, the library's
    # system, hierarchy and solver workspace included) after the timed steps
    used = total - torch.cuda.mem_get_info()[0]
    # FP64 residuals of the last step's forward and adjoint solves (not timed)
    A = lap.A_ff()
    # k = 1 FP32 SpMV of the 2.46M-row matrix: 176 MB of matrix and vectors
    spmv = spmv_k1(lib, A, sp, reps=50)
    M = sps.csr_matrix((A.values, A.col_idx, A.row_ptr), shape=(n, n))
    Xh, Bh, Gh, Lh = X.cpu().numpy(), B.cpu().numpy(), G.cpu().numpy(), LAM.cpu().numpy()
    rf = float((np.linalg.norm(Bh - M @ Xh, axis=0) / np.linalg.norm(Bh, axis=0)).max())
    ra = float((np.linalg.norm(Gh - M @ Lh, axis=0) / np.linalg.norm(Gh, axis=0)).max())
    line = {
        "metric": "poses/s, forward + adjoint at 1e-9 (2.6M-vert icosphere, 64 loops, 3 RHS)",
        "value": P * args.steps / el,
        "unit": "poses/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * el / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 inner / f64 refinement" if args.precision == "mixed" else "f64",
        "data": "synthetic (seeded icosphere, seeded constraint displacements and targets)",
        "config": {"workload": f"icosphere level 9 (nu 512), {mesh.nv} verts, 64 tilted loops, "
                               f"n_free={n}, n_con={ncon}, nnz={lap.info.nnz_ff}, {P} poses "
                               f"x 3 RHS, HSC theta {args.theta} ({hsc.info.nlevels} levels, "
                               f"coarse {hsc.info.n_coarse}), tol {args.tol}"},
        "spmv_k1": spmv,
        "tune": tune,
        "device_memory": {"in_use_after_steps_bytes": int(used), "total_bytes": int(total),
                          "in_use_before_steps_bytes": int(total - free0)},
        "solve": {"forward_iters": [f for f, _ in its], "adjoint_iters": [a for _, a in its],
                  "relres_forward": rf, "relres_adjoint": ra,
                  "setup_s": time.time() - t0 - el},
    }
    print(json.dumps(line))


def setup_bench(args):
    """The set-up path of a mesh, stage by stage: the seam cut (halfedges and
    seam duplication), the system build (halfedges, incoming lists, CSR
    pattern, assembly plan, values), the HSC hierarchy, the harmonic RHS
    table; then the re-assembly of new positions (cn_system_set_positions),
    kernel by kernel with HIP events, against its algorithmic bytes.  One
    JSON line per mesh."""
    import numpy as np
    import torch

    import tool_godot_curvenet_amd as cn
    from tool_godot_curvenet_amd import _lib, meshgen

    assert torch.cuda.is_available(), "no GPU"
    lib = _lib.load()
    dev = torch.device("cuda:0")
    peak = 8000.0
    for name, reps in (("capsules100k", 20), ("icosphere9", 10)):
        t = time.perf_counter()
        mesh = meshgen.config_mesh(name)
        t_mesh = time.perf_counter() - t
        print(f"setup {name}: mesh generated in {t_mesh:.1f} s", file=sys.stderr, flush=True)

        def timed(fn):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            r = fn()
            torch.cuda.synchronize()
            return r, time.perf_counter() - t0

        (nt, origin, con, nv2), t_cut = timed(
            lambda: cn.cut_seams(mesh.nv, mesh.tris, mesh.cut_edges))
        pos = np.ascontiguousarray(mesh.pos[origin])
        lap, t_sys = timed(lambda: cn.CutLaplacian(pos, nt, con))
        hsc, t_hsc = timed(lambda: cn.HSC(lap, theta=args.theta))
        n, ncon, nv, nf = lap.n_free, lap.n_con, lap.info.nv, lap.info.nf
        nnz_ff, nnz_fc = lap.info.nnz_ff, lap.info.nnz_fc
        rng = np.random.default_rng(meshgen.BASE_SEED + 60)
        Xc = torch.as_tensor(rng.uniform(-0.2, 0.2, size=(ncon, 3)), device=dev)
        _, t_rhs_first = timed(lambda: lap.harmonic_rhs(Xc, as_numpy=False))
        _, t_rhs = timed(lambda: lap.harmonic_rhs(Xc, as_numpy=False))
        print(f"setup {name}: cut {t_cut:.2f} s, system {t_sys:.2f} s, hierarchy "
              f"{t_hsc:.2f} s", file=sys.stderr, flush=True)
        # re-assembly of the same positions: the values must not change
        a0 = lap.A_ff().values.copy()
        pos_d = torch.as_tensor(pos, device=dev)
        import ctypes as C
        sp = C.c_void_p(torch.cuda.current_stream().cuda_stream)
        names = ("asm_face", "asm_values", "asm_diag")
        ms = {k: 0.0 for k in names}
        lib.cn_profile_enable(1)
        wall = []
        for _ in range(reps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            _lib.check(lib.cn_system_set_positions(lap.handle, C.c_void_p(pos_d.data_ptr()),
                                                   sp), "set_positions")
            wall.append(time.perf_counter() - t0)
            for k in names:
                v = C.c_double(0.0)
                lib.cn_profile_query(k.encode(), C.byref(v))
                ms[k] += v.value
        lib.cn_profile_enable(0)
        same = bool(np.array_equal(lap.A_ff().values, a0))
        # terms of the assembly plan: 3 per face corner at a free vertex
        nterms = int(3 * np.count_nonzero(con[nt.reshape(-1)] == 0))
        tot = nnz_ff + nnz_fc
        by = {"asm_face": 12 * nf + 24 * nv + 72 * nf,
              "asm_values": 4 * (tot + 1) + 4 * nterms + 72 * nf + 12 * tot,
              "asm_diag": 20 * n}
        kern = {}
        for k in names:
            avg = ms[k] / reps
            gbs = by[k] / (avg * 1e-3) / 1e9 if avg > 0 else 0.0
            kern[k] = {"avg_ms": avg, "bytes_per_launch": by[k], "achieved_gbs": gbs,
                       "frac": gbs / peak}
        line = {
            "metric": f"set-up stages and re-assembly ({name})",
            "workload": name,
            "nv": nv, "nf": nf, "n_free": n, "n_con": ncon, "nnz_ff": nnz_ff,
            "nnz_fc": nnz_fc, "assembly_terms": nterms,
            "stages_s": {"mesh_generation_host": t_mesh, "cut_seams": t_cut,
                         "system_build": t_sys, "hsc_build": t_hsc,
                         "hsc_levels": hsc.info.nlevels, "hsc_coarse": hsc.info.n_coarse,
                         "harmonic_rhs_first_call": t_rhs_first, "harmonic_rhs": t_rhs},
            "reassembly": {"reps": reps, "wall_ms_median": 1e3 * float(np.median(wall)),
                           "kernels": kern, "values_unchanged": same},
            "device": torch.cuda.get_device_name(0),
        }
        print(json.dumps(line), flush=True)
        del lap, hsc, pos_d, Xc
        torch.cuda.empty_cache()


def ddm_bench(args):
    """Direct Delta Mush on the 100k capsule union: the bind (unit-perturbation
    harvest of every handle with the HSC-preconditioned multi-RHS PCG, two
    smoothing passes, top-8 sparsification), then the runtime skinning matvec
    (FP32, 8 influences a vertex) over 256 poses of per-handle transforms --
    one step is one matvec launch over all poses."""
    import numpy as np
    import torch

    import oracle as O
    import tool_godot_curvenet_amd as cn
    from tool_godot_curvenet_amd import ddm, meshgen

    assert torch.cuda.is_available(), "no GPU"
    dev = torch.device("cuda:0")
    t0 = time.time()
    mesh = meshgen.config_mesh("capsules100k")
    lap = cn.CutLaplacian.from_mesh(mesh)
    hsc = cn.HSC(lap, theta=args.theta)
    nv, H, K, P = lap.info.nv, lap.n_con, 8, args.poses
    print(f"ddm: system and hierarchy at {time.time() - t0:.1f} s (nv {nv}, handles {H})",
          file=sys.stderr, flush=True)

    def bind_once():
        ts = [time.perf_counter()]
        W, infos = ddm.bind_harvest(lap, hsc=hsc, block=768, precision=args.precision,
                                    tol=args.tol)
        torch.cuda.synchronize()
        ts.append(time.perf_counter())
        ap, aj = ddm.vertex_adjacency(nv, lap.tris)
        Ws = ddm.smooth_weights(W, ap, aj, 2, 0.5)
        torch.cuda.synchronize()
        ts.append(time.perf_counter())
        del W
        idx, w = ddm.sparsify_topk(Ws, K)
        torch.cuda.synchronize()
        ts.append(time.perf_counter())
        pou = float((Ws.sum(dim=1) - 1.0).abs().max().item())
        return idx, w, infos, np.diff(ts), pou

    bind_once()                                         # warm-up (allocations, code load)
    idx, w, infos, tb, pou = bind_once()
    print(f"ddm: bind {tb.sum():.3f} s (harvest {tb[0]:.3f}, smooth {tb[1]:.3f}, "
          f"top-K {tb[2]:.3f})", file=sys.stderr, flush=True)

    # the bind-cache sidecar: hierarchy + influences written once, read back
    # by a later session instead of rebuilding
    import tempfile
    from tool_godot_curvenet_amd import formats
    th = time.perf_counter()
    hsc2 = cn.HSC(lap, theta=args.theta)
    torch.cuda.synchronize()
    t_build = time.perf_counter() - th
    start0 = torch.arange(0, (nv + 1) * K, K, dtype=torch.int32, device=dev)
    with tempfile.TemporaryDirectory() as td:
        side = os.path.join(td, "bind.safetensors")
        tw = time.perf_counter()
        formats.save_sidecar(side, lap, hsc=hsc2, influences=(start0, idx.reshape(-1), w.reshape(-1)))
        t_write = time.perf_counter() - tw
        side_bytes = os.path.getsize(side)
        tl = time.perf_counter()
        hsc3 = formats.load_hsc(side, lap)
        formats.load_influences(side, lap)
        torch.cuda.synchronize()
        t_load = time.perf_counter() - tl
    r_chk = torch.randn(lap.n_free, 3, dtype=torch.float64, device=dev)
    same = bool(np.array_equal(hsc2.apply(r_chk), hsc3.apply(r_chk)))
    del hsc2, hsc3
    print(f"ddm: sidecar {side_bytes / 1e6:.1f} MB, write {t_write:.3f} s, load {t_load:.3f} s, "
          f"HSC build {t_build:.3f} s", file=sys.stderr, flush=True)

    # runtime: per-handle rigid transforms for P poses (FP32)
    rng = np.random.default_rng(meshgen.BASE_SEED + 60)
    ang = rng.uniform(-0.3, 0.3, size=(P, H, 3))
    T = np.zeros((P, H, 4, 4))
    for a in range(3):
        c, s_ = np.cos(ang[..., a]), np.sin(ang[..., a])
        R = np.zeros((P, H, 3, 3))
        i, j = [(1, 2), (0, 2), (0, 1)][a]
        R[..., a, a] = 1.0
        R[..., i, i], R[..., j, j], R[..., i, j], R[..., j, i] = c, c, -s_, s_
        T[..., :3, :3] = R if a == 0 else T[..., :3, :3] @ R
    T[..., :3, 3] = rng.normal(0, 0.05, size=(P, H, 3))
    T[..., 3, 3] = 1.0
    Tg = torch.as_tensor(T.reshape(P, H, 16), dtype=torch.float32, device=dev)
    rest = torch.as_tensor(lap.pos, dtype=torch.float32, device=dev)
    start = torch.arange(0, (nv + 1) * K, K, dtype=torch.int32, device=dev)
    idxg = idx.reshape(-1).contiguous()
    wg = w.reshape(-1).to(torch.float32).contiguous()
    out = torch.empty((P, nv, 3), dtype=torch.float32, device=dev)
    # distinct handles of each 256-vertex tile (the kernel stages them in LDS)
    ih = idxg.cpu().numpy().reshape(nv, K)
    tile_nu = np.array([len(np.unique(ih[i:i + 256])) for i in range(0, nv, 256)])
    from tool_godot_curvenet_amd import _lib
    from tool_godot_curvenet_amd.curvenet import _p, _stream
    lib = _lib.load()

    def step():
        _lib.check(lib.cn_lbs_matvec_f32(nv, H, P, _p(rest), _p(Tg), _p(start), _p(idxg),
                                         _p(wg), _p(out), _stream()), "lbs_matvec_f32")

    for _ in range(args.warmup):
        step()
    steps = args.steps * 20                             # one launch is ~0.1 ms
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t1 = time.perf_counter()
    e0.record()
    for _ in range(steps):
        step()
    e1.record()
    torch.cuda.synchronize()
    el = time.perf_counter() - t1
    ms = e0.elapsed_time(e1) / steps
    # algorithmic bytes of one launch: rest, row pointers, influences, output,
    # and every pose's transform table once
    alg = nv * (12 + 4 + 8 * K) + P * nv * 12 + P * H * 64
    # check one pose against the oracle (FP64 on the FP32 inputs)
    p = P // 3
    ref = O.lbs_matvec(lap.pos.astype(np.float32).astype(np.float64),
                       T.reshape(P, H, 16)[p].astype(np.float32).astype(np.float64),
                       start.cpu().numpy(), idxg.cpu().numpy(),
                       wg.cpu().numpy().astype(np.float64))
    err = float(np.abs(out[p].cpu().numpy() - ref).max())
    cpu = None
    if not args.no_cpu:
        cp = 32
        tc = time.perf_counter()
        for q in range(cp):
            O.lbs_matvec(lap.pos, T.reshape(P, H, 16)[q], start.cpu().numpy(),
                         idxg.cpu().numpy(), w.reshape(-1).cpu().numpy())
        tcs = time.perf_counter() - tc
        cpu = {"value": cp * nv / tcs, "unit": "vertex-poses/s", "cores": 1, "kind": "port",
               "sample": f"oracle lbs_matvec (FP64, one thread) over {cp} poses of the "
                         f"same influences and transforms ({tcs:.2f} s)",
               "cpu": _cpu_model()}
    line = {
        "metric": "vertex-poses/s, DDM skinning matvec (100k capsule union, 8 influences)",
        "value": P * nv / (ms * 1e-3),
        "unit": "vertex-poses/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (capsule union, harvested weights, seeded per-handle rigid "
                "transforms)",
        "config": {"workload": f"capsule union {nv} verts, {H} handles, top-{K} influences, "
                               f"{P} poses; bind: harvest (HSC PCG, {args.precision}, tol "
                               f"{args.tol}, blocks of 768 handles), 2 smoothing passes "
                               f"(omega 0.5), top-{K}"},
        "roofline": {"bound": "hbm", "achieved": alg / (ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "traffic": None, "kernel": "k_lbs<float>", "avg_ms": ms,
                     "bytes_per_launch": alg},
        "cpu_baseline": cpu,
        "bind": {"seconds": float(tb.sum()), "harvest_s": float(tb[0]),
                 "smooth_s": float(tb[1]), "topk_s": float(tb[2]),
                 "handles_per_s": H / float(tb[0]),
                 "iters": [i.iters for i in infos], "partition_of_unity_err": pou},
        "check": {"pose": p, "max_abs_err_vs_oracle": err},
        "sidecar": {"bytes": side_bytes, "write_s": t_write, "load_s": t_load,
                    "hsc_build_s": t_build, "apply_identical": same},
        "tiles": {"count": int(len(tile_nu)), "handles_mean": float(tile_nu.mean()),
                  "handles_max": int(tile_nu.max())},
        "wall_s": el,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
