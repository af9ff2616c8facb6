"""Autotuner of the solver's launch shapes (cn_tune), with a JSON cache.

The fused PCG's kernels have a few launch-shape choices whose best value
depends on the matrix (rows, nonzeros, bandwidth), the number of right-hand
sides k and the precision:

* ``spmm_form``: the LDS-staged SpMM (a row range of one column tile per
  block, neighbours read from an LDS window) or the row-table SpMM (one
  wavefront per row), and for the latter ``spmm_gens``, the row stripes
  each XCD's share is walked in;
* ``level_cap``: the grid cap of the HSC level passes (a row is a chain of
  dependent loads: more waves in flight hide more latency, up to the point
  where they only add scheduling);
* ``tail_rows``: the size below which HSC levels run in the fused tail
  launch (one workgroup per column slice walks them with barriers);
* ``chain_levels``: the levels one group of down-sweep launches chains
  through (each row is written once a group; longer groups write less and
  gather rows of lower levels again), and ``down_cap``, the down sweep's own
  grid cap;
* ``coarse_prec``: the HSC coarse product of the mixed-precision apply on
  the FP64 inverse, its FP32 rounding, or (wide products) its split into
  two bfloat16 matrices.

With mixed precision it also picks the solve's ``inner_tol``, the factor by
which a column's iterated residual falls between two residual replacements
(FP64 x update and recomputed b - A x): a larger drop saves replacements, a
too large one lets the FP32 recurrence drift from the true residual and
costs iterations.  The choice is kept on the system (``system.inner_tol``),
which its solves use unless the caller passes another.

``autotune`` times full solves of the caller's system and right-hand side
for a short list of candidates (coordinate search: the SpMM form first, then
the level cap, the tail size, the chain length, inner_tol), keeps the fastest, and records it in a
JSON file keyed by device, rows, nonzeros, k, precision and preconditioner.
``lookup`` reads that file, ``apply`` sets a choice on a system.  The shapes
change no result beyond rounding (the SpMM forms compute the same products;
only partial sums are grouped differently), and every candidate's solve is
checked to reach the tolerance.

The cache lives in ``$CN_AUTOTUNE_CACHE`` if set, else
``~/.cache/tool_godot_curvenet_amd/autotune.json`` (the system temp
directory when the home directory is not writable).
"""
from __future__ import annotations

import json
import os
import tempfile
import time

from . import _lib as L

FIELDS = ("spmm_form", "spmm_gens", "level_cap", "tail_rows", "chain_levels", "coarse_prec",
          "down_cap")
DEFAULT = {"spmm_form": 0, "spmm_gens": 0, "level_cap": 0, "tail_rows": 0, "chain_levels": 0,
           "coarse_prec": 0, "down_cap": 0}
INNER_TOLS = (0.0316, 0.01, 1e-3, 1e-4)   # the first is the built-in one


def cache_path() -> str:
    p = os.environ.get("CN_AUTOTUNE_CACHE")
    if p:
        return p
    home = os.path.join(os.path.expanduser("~"), ".cache", "tool_godot_curvenet_amd")
    try:
        os.makedirs(home, exist_ok=True)
        if os.access(home, os.W_OK):
            return os.path.join(home, "autotune.json")
    except OSError:
        pass
    return os.path.join(tempfile.gettempdir(), "tool_godot_curvenet_amd_autotune.json")


def _load(path: str) -> dict:
    try:
        with open(path) as fh:
            d = json.load(fh)
        return d if isinstance(d, dict) else {}
    except (OSError, ValueError):
        return {}


def _store(path: str, key: str, rec: dict) -> None:
    d = _load(path)
    d[key] = rec
    tmp = f"{path}.{os.getpid()}.tmp"
    try:
        with open(tmp, "w") as fh:
            json.dump(d, fh, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        try:
            os.remove(tmp)
        except OSError:
            pass


def _device() -> str:
    import torch
    p = torch.cuda.get_device_properties(0)
    return f"{p.name}/{getattr(p, 'gcnArchName', '')}/{p.multi_processor_count}cu"


def key_of(system, k: int, precision: str, precond: str) -> str:
    i = system.info
    return (f"{_device()}|n={i.n_free}|nnz={i.nnz_ff}|nseg={getattr(system, 'nseg', 1)}|k={k}"
            f"|{precision}|{precond}")


def lookup(system, k: int, precision: str = "mixed", precond: str = "hsc",
           path: str | None = None) -> dict | None:
    """The cached choice for this system shape, or None."""
    rec = _load(path or cache_path()).get(key_of(system, k, precision, precond))
    if not rec or not isinstance(rec.get("tune"), dict):
        return None
    return {f: int(rec["tune"].get(f, 0)) for f in FIELDS}


def apply(system, tune: dict | None) -> None:
    system.set_tune(tune)


def _tail_applies(k: int, precision: str) -> bool:
    vec = (4 if k % 4 == 0 else (2 if k % 2 == 0 else 1)) if precision == "mixed" else \
        (2 if k % 2 == 0 else 1)
    return k > 4 and k // vec >= 128


def autotune(system, B, precond: str = "hsc", precision: str = "mixed", tol: float = 1e-9,
             reps: int = 2, path: str | None = None, use_cache: bool = True,
             shapes: bool = True, **pre) -> dict:
    """Time full solves of system on right-hand side B (n x k, FP64) for the
    candidate launch shapes (and, in mixed precision, residual-replacement
    drops) and keep the fastest (set on the system and stored in the
    cache).  `pre` passes the preconditioner object (hsc=, ic0= or sgs=).
    shapes=False keeps the system's launch shapes and picks only the drop
    (for another right-hand side of the same system, such as an adjoint's;
    nothing is cached then).
    Returns {"key", "tune", "inner_tol", "ms", "cached", "candidates"}."""
    import torch
    path = path or cache_path()
    k = int(B.shape[1]) if len(B.shape) == 2 else 1
    key = key_of(system, k, precision, precond)
    if use_cache and shapes:
        hit = lookup(system, k, precision, precond, path)
        if hit is not None:
            system.set_tune(hit)
            rec = _load(path)[key]
            system.inner_tol = float(rec.get("inner_tol", INNER_TOLS[0]))
            return {"key": key, "tune": hit, "inner_tol": system.inner_tol, "ms": rec.get("ms"),
                    "cached": True, "candidates": rec.get("candidates", [])}
    tried = []

    def run(tune: dict, inner_tol: float = INNER_TOLS[0], n: int = reps) -> float:
        system.set_tune(tune)
        kw = dict(precond=precond, precision=precision, tol=tol, inner_tol=inner_tol, **pre)
        system.solve(B, **kw)  # warm-up
        best = float("inf")
        for _ in range(n):
            torch.cuda.synchronize()
            t = time.perf_counter()
            _, info = system.solve(B, **kw)
            torch.cuda.synchronize()
            el = time.perf_counter() - t
            if info.status != L.CN_OK:
                return float("inf")     # a shape that does not converge is not a candidate
            best = min(best, el)
        tried.append({"tune": dict(tune), "inner_tol": inner_tol, "ms": 1e3 * best})
        return best

    if not shapes:
        cur = system.get_tune()
        best, itol = float("inf"), INNER_TOLS[0]
        if precision == "mixed":
            for it in INNER_TOLS:
                t = run(cur, it, max(reps, 4))
                if t < best:
                    best, itol = t, it
        system.inner_tol = itol
        return {"key": key, "tune": cur, "inner_tol": itol, "ms": 1e3 * best, "cached": False,
                "candidates": tried}
    cur = dict(DEFAULT)
    # 1: the SpMM form (LDS-staged, or the row table with 4, 2 or 1 stripes)
    opts = [{"spmm_form": 2, "spmm_gens": 0}] + [{"spmm_form": 1, "spmm_gens": g}
                                                 for g in (4, 2, 1)]
    best_t, best = None, float("inf")
    for o in opts:
        t = run({**cur, **o})
        if t < best:
            best, best_t = t, o
    if best_t is not None:
        cur.update(best_t)
    # 2: the HSC level-pass grid cap
    if precond == "hsc":
        for cap in (2048, 4096, 65536):
            t = run({**cur, "level_cap": cap})
            if t < best:
                best, cur = t, {**cur, "level_cap": cap}
        # 3: the fused tail's level size (only where the tail runs)
        if _tail_applies(k, precision):
            for rows in (-1, 4096):
                t = run({**cur, "tail_rows": rows})
                if t < best:
                    best, cur = t, {**cur, "tail_rows": rows}
        # 4: the down sweep's chain length (the wide form, k > 4)
        if k > 4:
            for g in (2, 4, 5, 6, 8):
                t = run({**cur, "chain_levels": g})
                if t < best:
                    best, cur = t, {**cur, "chain_levels": g}
            # the down sweep's own grid cap
            for cap in (16384, 65536):
                t = run({**cur, "down_cap": cap})
                if t < best:
                    best, cur = t, {**cur, "down_cap": cap}
        # the coarse product on the FP32 rounding of the inverse (mixed
        # precision): twice the matrix rate (k > 4, MFMA) or half the bytes
        # (k <= 4), against what the perturbed preconditioner costs in
        # iterations
        # (k > 4 also the split-bfloat16 product, 3: three bfloat16 MFMAs a
        # product at eight times the FP32 matrix rate each, a b to 2^-16)
        if precision == "mixed":
            for cp in ((2, 3) if k > 4 else (2,)):
                t = run({**cur, "coarse_prec": cp})
                if t < best:
                    best, cur = t, {**cur, "coarse_prec": cp}
    # 5: the residual-replacement drop (mixed precision); the candidates
    # differ by a replacement or two (a few per cent), so each, the built-in
    # one included, is timed again over more solves
    itol = INNER_TOLS[0]
    if precision == "mixed":
        best = float("inf")
        for it in INNER_TOLS:
            t = run(cur, it, max(reps, 4))
            if t < best:
                best, itol = t, it
    system.set_tune(cur)
    system.inner_tol = itol
    rec = {"tune": cur, "inner_tol": itol, "ms": 1e3 * best, "candidates": tried,
           "when": time.strftime("%Y-%m-%dT%H:%M:%S")}
    _store(path, key, rec)
    return {"key": key, "tune": cur, "inner_tol": itol, "ms": 1e3 * best, "cached": False,
            "candidates": tried}
