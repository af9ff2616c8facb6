"""Build the gfx950 shared library ``libcurvenet_la.so`` in-tree.

Every ``csrc/*.hip`` file is compiled with hipcc for gfx950 only (host-only
``csrc/*.cpp`` files with g++) and linked into one shared object that exports
the C ABI of ``include/curvenet_la.h``.

Flags:
  * ``-Xclang -target-feature -Xclang -packed-fp32-ops`` keeps the compiler
    from emitting packed FP32 instructions;
  * the dense FP64 kernels are built with ``-mllvm -amdgpu-mfma-vgpr-form``
    (MFMA accumulators in VGPRs);
  * files whose FP64 results must match the spec bit for bit (topology,
    assembly, IC(0), SpMV) are built with ``-ffp-contract=off``; the solver
    file keeps contraction on (its FP32 arithmetic uses explicit fma()).

Run ``python -m tool_godot_curvenet_amd.build`` or call :func:`build`.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import pathlib
import shutil
import subprocess
import sys

HERE = pathlib.Path(__file__).resolve().parent
CSRC = HERE / "csrc"
LIB = HERE / "libcurvenet_la.so"
OBJ = HERE / "_build"

EXACT = {"cn_topology.hip", "cn_assembly.hip", "cn_ic0.hip", "cn_spmv.hip", "cn_hsc_build.hip", "cn_hsc_host.cpp", "cn_sgs.hip", "cn_deform.hip", "cn_ddm.hip"}
# files whose MFMA accumulators stay in VGPRs (with the AGPR form the
# compiler copies loop-carried accumulators between the two files each trip)
MFMA_VGPR = {"cn_dense.hip"}


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _flags(name: str) -> list[str]:
    f = [
        "--offload-arch=gfx950",
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-Wall",
        "-Wno-unused-function",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        f"-I{HERE.parent / 'include'}",
    ]
    if name in EXACT:
        f.append("-ffp-contract=off")
    if name in MFMA_VGPR:
        f += ["-mllvm", "-amdgpu-mfma-vgpr-form"]
    return f


def _compile(src: pathlib.Path) -> pathlib.Path:
    out = OBJ / (src.stem + ".o")
    if src.suffix == ".cpp":  # host-only code
        cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-pthread", "-Wall",
               f"-I{HERE.parent / 'include'}", "-c", str(src), "-o", str(out)]
        if src.name in EXACT:
            cmd.insert(1, "-ffp-contract=off")
    else:
        cmd = [_hipcc(), *_flags(src.name), "-c", str(src), "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src.name}:\n{r.stderr}")
    # host-pass warnings about the unknown target feature are expected
    return out


def build(verbose: bool = False) -> pathlib.Path:
    OBJ.mkdir(exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp"))
    jobs = min(len(srcs), int(os.environ.get("MAX_JOBS", "8")), 8)
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    # --no-undefined: a declaration that matches no definition fails here,
    # not at load time
    cmd = [_hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-pthread", "-Wl,--no-undefined",
           "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
