"""Per-level launch-shape autotuner.

For every level of a problem and every tunable operator (SpMV, the four
SymGS sweep kernels on their wide and on their narrow colours, the fused
restriction) and both precisions, time each
compiled launch shape -- rows per thread {1, 2} x threads per block
{64, 128, 256} -- with HIP events on the solver stream and keep the fastest.
Results are cached as JSON keyed by the level's grid and the precision, so a
later run on the same grids applies them without timing.  `tune_batch` does the
same for a batch of problems, per level, from short batched solves.  The
reference has no tuner; it hard-codes 1024/256/128-thread blocks
(ComputeSPMV.cpp:262, ComputeSYMGS.cpp:395, ComputeRestriction.cpp:214).
"""
from __future__ import annotations

import json
import os

from . import (OP_BWD_FULL, OP_BWD_ZERO, OP_FWD_FULL, OP_FWD_ZERO, OP_RESTRICT, OP_SPMV,
               SH_BWD_FULL, SH_BWD_FULL_N, SH_BWD_ZERO, SH_BWD_ZERO_N, SH_FWD_FULL, SH_FWD_FULL_N,
               SH_FWD_ZERO, SH_FWD_ZERO_N, SH_RESTRICT, SH_SPMV, STORE_F64, STORE_MIXED)

SHAPES = [(block, rows) for rows in (1, 2) for block in (64, 128, 256)]
OPS = {  # tunable operator -> timing operator (a sweep's wide and narrow
    # colours are timed together; the other set keeps its current shape)
    SH_SPMV: OP_SPMV, SH_FWD_ZERO: OP_FWD_ZERO, SH_BWD_ZERO: OP_BWD_ZERO,
    SH_FWD_FULL: OP_FWD_FULL, SH_BWD_FULL: OP_BWD_FULL, SH_RESTRICT: OP_RESTRICT,
    SH_FWD_ZERO_N: OP_FWD_ZERO, SH_BWD_ZERO_N: OP_BWD_ZERO, SH_FWD_FULL_N: OP_FWD_FULL,
    SH_BWD_FULL_N: OP_BWD_FULL,
}
OP_NAMES = {SH_SPMV: "spmv", SH_FWD_ZERO: "fwd_zero", SH_BWD_ZERO: "bwd_zero",
            SH_FWD_FULL: "fwd", SH_BWD_FULL: "bwd", SH_RESTRICT: "restrict",
            SH_FWD_ZERO_N: "fwd_zero_narrow", SH_BWD_ZERO_N: "bwd_zero_narrow", SH_FWD_FULL_N: "fwd_narrow",
            SH_BWD_FULL_N: "bwd_narrow"}
DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "rochpcg_amd", "shapes.json")


def _key(level, mixed):
    return f"{level.nx}x{level.ny}x{level.nz}:{'mixed' if mixed else 'f64'}"


def load(path=DEFAULT_CACHE):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def save(table, path=DEFAULT_CACHE):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def apply(prob, table):
    """Apply cached shapes to every level/precision found in `table`; returns count."""
    n = 0
    for lv, L in enumerate(prob.levels):
        for mixed in (False, True):
            if not (L.stores & (STORE_MIXED if mixed else STORE_F64)):
                continue
            entry = table.get(_key(L, mixed))
            if not entry:
                continue
            for op, name in OP_NAMES.items():
                if name in entry:
                    block, rows = entry[name]["shape"]
                    prob.set_shape(block, rows, level=lv, op=op, mixed=mixed)
                    n += 1
    return n


def _prime(prob, mixed):
    """Operands as in a solve: r = b - A*0 from the CG preamble, and every
    level's r / x filled by one V-cycle."""
    from . import CG_F64, CG_MIXED, OP_MG
    prob.cg_begin(mode=CG_MIXED if mixed else CG_F64, max_iter=1, tol=0.0)
    prob.cg_end()
    prob.time_operator(OP_MG, level=0, mixed=mixed, reps=1)


def tune(prob, mixed=False, reps=3, rounds=3, levels=None, table=None, log=None):
    """Time every shape of every tunable operator -- candidates interleaved
    over `rounds` rounds, median per candidate -- then apply the fastest and
    return the table."""
    import statistics
    table = {} if table is None else table
    last = prob.nlevels - 1
    _prime(prob, mixed)
    for lv in (range(prob.nlevels) if levels is None else levels):
        L = prob.levels[lv]
        if not (L.stores & (STORE_MIXED if mixed else STORE_F64)):
            continue
        entry = table.setdefault(_key(L, mixed), {})
        for op, top in OPS.items():
            if op == SH_RESTRICT and lv == last:
                continue
            cands = SHAPES if op != SH_RESTRICT else [(b, 1) for b in (64, 128, 256)]
            samples = {c: [] for c in cands}
            for _ in range(rounds):
                for block, rows in cands:
                    prob.set_shape(block, rows, level=lv, op=op, mixed=mixed)
                    samples[(block, rows)].append(prob.time_operator(top, level=lv, mixed=mixed, reps=reps))
            timings = {c: statistics.median(v) for c, v in samples.items()}
            (block, rows), best = min(timings.items(), key=lambda kv: kv[1])
            prob.set_shape(block, rows, level=lv, op=op, mixed=mixed)
            default = timings.get((256, 1))
            entry[OP_NAMES[op]] = {"shape": [block, rows], "ms": round(best, 5),
                                   "default_ms": None if default is None else round(default, 5)}
            if log:
                log(f"tune {_key(L, mixed)} {OP_NAMES[op]:9s} best block={block} rows={rows} "
                    f"{best * 1e3:9.1f} us (256x1: {default * 1e3:9.1f} us)")
    return table


def _batch_seconds(B, mode, iters):
    """Seconds of `iters` graph-replayed batched CG iterations from x = 0 (the
    graph captured beforehand)."""
    import time

    import numpy as np

    for P in B.problems:
        P.vec["x"].upload(np.zeros(P.m))
    B.cg(max_iter=2, mode=mode)
    for P in B.problems:
        P.vec["x"].upload(np.zeros(P.m))
    t = time.perf_counter()
    B.cg(max_iter=iters, mode=mode)
    return time.perf_counter() - t


def tune_batch(B, nlevels, modes, iters=10, blocks=(64, 128, 256), rows=(1, 2), log=None):
    """Launch shape of the batched kernels, per level: threads per block (all
    kernels) and rows per thread (sweeps).  The best uniform shape first, then
    each level in turn keeps the shape that minimises the summed time of
    `iters` iterations in every mode of `modes`.  Returns the per-level
    (block, rows) pairs (already applied)."""
    shapes = [(b, r) for r in rows for b in blocks]

    def cost():
        return sum(_batch_seconds(B, m, iters) for m in modes)

    def apply(per_level):
        for lv, (b, r) in enumerate(per_level):
            B.set_block(lv, b)
            B.set_rows(lv, r)

    best = None
    for sh in shapes:
        apply([sh] * nlevels)
        t = cost()
        if best is None or t < best[0]:
            best = (t, [sh] * nlevels)
    for lv in range(nlevels):
        for sh in shapes:
            if sh == best[1][lv]:
                continue
            trial = list(best[1])
            trial[lv] = sh
            apply(trial)
            t = cost()
            if t < best[0]:
                best = (t, trial)
    apply(best[1])
    if log:
        log(f"batch (block, rows) per level {best[1]} ({best[0] * 1e3:.1f} ms for {iters} iterations per mode)")
    return best[1]
