#!/usr/bin/env python3
"""HPCG solver benchmark on one MI355X.

A step is one preconditioned CG iteration (4-level MG V-cycle with one pre-
and one post-smoothing multicolour SymGS per level, outer SpMV, dots and
vector updates) on the 27-point problem, all arrays resident in HBM.

Prints one JSON line: the FP64 HPCG-formula GFLOP/s (value), the mixed-
precision CG's GFLOP/s and its time to a relative residual of 1e-9 against
the FP64 CG's, the modelled HBM TB/s of both, a roofline entry for the outer
SpMV kernel timed with HIP events on the solver stream, and the CPU baseline
(the oracle's CG_ref restatement, OpenMP, on this host's cores).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "HPCG GFLOP/s + HBM TB/s: mixed vs FP64 CG to rel. residual 1e-9, 1 GPU"
HBM_PEAK_GBPS = 8000.0


def profiled_traffic(dims):
    """HBM bytes per launch of the fused SpMV + p'Ap kernel (FETCH_SIZE +
    WRITE_SIZE, copy-calibrated) from the committed rocprofv3 summary of this
    grid, or None when there is none."""
    import re
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{dims[0]}x{dims[1]}x{dims[2]}.md")
    try:
        with open(path) as f:
            for line in f:
                m = re.match(r"^roofline traffic \(k_spmv, HBM bytes per launch\): (\d+)", line)
                if m:
                    return int(m.group(1))
    except OSError:
        pass
    return None


def spmv_bytes(m, vb=8):
    return m * (27 * (4 + vb) + 8 + 8)


def cpu_baseline(dims, iters, threads):
    """CG_ref on the host, natural order, b = A*ones, x0 = 0: the reference's own
    _ref code (CG_ref.cpp, ComputeMG_ref.cpp, ComputeSYMGS_ref.cpp, ... built
    into oracle/_ref) when that library is present, else the oracle's
    restatement of it.  OpenMP where the reference has pragmas; SymGS serial."""
    import oracle
    oracle.set_threads(threads)
    t0 = time.perf_counter()
    P = oracle.Problem(*dims, nlevels=4, order=oracle.ORDER_NATURAL)
    b = P.rhs_ones()
    R = oracle.RefHierarchy(P) if oracle.ref_available() else None
    setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    res = R.cg(b, max_iter=iters, tol=0.0) if R else P.cg(b, max_iter=iters, tol=0.0)
    el = time.perf_counter() - t0
    flops = oracle.hpcg_flops_per_iter(dims) * res["niters"]
    what = ("the reference's own CG_ref and _ref kernels (built from its sources into oracle/_ref)"
            if R else "oracle CG_ref restatement")
    if R:
        R.close()
    P.close()
    return dict(value=round(flops / el / 1e9, 3), unit="GFLOP/s", cores=threads,
                kind="reference" if R else "port",
                sample=f"{what}, natural order, OpenMP except SymGS, at "
                       f"{dims[0]}x{dims[1]}x{dims[2]}, {res['niters']} iterations, "
                       f"{el:.1f} s timed ({setup:.1f} s setup untimed)",
                s_per_iter=round(el / max(res["niters"], 1), 3))


def batched_leg(hip, count, log):
    """64 independent problems, nx/ny/nz each from {16,...,96} (splitmix64 seed 7),
    problem k with x* uniform in [-1,1] from seed 7+k; solved to 1e-9 together
    (one launch per operator/level/colour for all problems, graph-replayed)."""
    import numpy as np
    from rochpcg_amd import vv
    dims = vv.batched_dims(count, seed=7)
    t0 = time.perf_counter()
    B = hip.Batch(dims, nlevels=4, stores=hip.STORE_F64 | hip.STORE_MIXED)
    setup = time.perf_counter() - t0
    for k, P in enumerate(B.problems):
        P.make_rhs(1, 7 + k)
    flops = np.array([vv.hpcg_flops_per_iteration(d) for d in dims])
    # threads per block of the batched kernels, per level (both CG modes)
    from rochpcg_amd import autotune
    blocks = autotune.tune_batch(B, 4, (hip.CG_F64, hip.CG_MIXED), log=log)
    res = {"problems": count, "rows_total": int(sum(a * b * c for a, b, c in dims)),
           "setup_s": round(setup, 3), "block": blocks}
    for name, mode in (("f64", hip.CG_F64), ("mixed", hip.CG_MIXED)):
        for P in B.problems:
            P.vec["x"].upload(np.zeros(P.m))
        B.cg(max_iter=2, mode=mode)   # capture this mode's graph
        for P in B.problems:
            P.vec["x"].upload(np.zeros(P.m))
        t = time.perf_counter()
        r = B.cg(max_iter=500, tol=1e-9, mode=mode)
        el = time.perf_counter() - t
        worst = float(np.max(r["normr"] / r["normr0"]))
        res[name] = {"seconds": round(el, 4), "iterations_max": int(r["niters"].max()),
                     "iterations_min": int(r["niters"].min()), "worst_rel_residual": float(f"{worst:.3e}"),
                     "gflops": round(float((flops * r["niters"]).sum()) / el / 1e9, 2)}
        log(f"batched {name}: {res[name]}")
    res["speedup_time_to_solution"] = round(res["f64"]["seconds"] / res["mixed"]["seconds"], 3)
    B.close()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--nx", type=int, default=320)
    ap.add_argument("--ny", type=int, default=None)
    ap.add_argument("--nz", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-solve", action="store_true", help="skip the solve-to-1e-9 comparison")
    ap.add_argument("--cpu-iters", type=int, default=5)
    ap.add_argument("--eager", action="store_true", help="launch kernels one by one instead of a hipGraph")
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch shapes (256 threads, 1 row)")
    ap.add_argument("--seed", type=int, default=1234, help="splitmix64 seed of x* in the solve-to-1e-9 leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched uneven-problems leg")
    ap.add_argument("--batch-count", type=int, default=64)
    ap.add_argument("--shapes", default=None,
                    help="apply launch shapes saved by --save-shapes instead of tuning")
    ap.add_argument("--save-shapes", default=None, help="write the tuned launch shapes to this JSON file")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; multi-GPU runs (--gpus > 1) "
                                   "are not part of it", "n_gpus": args.gpus}))
        sys.exit(2)
    dims = (args.nx, args.ny or args.nx, args.nz or args.nx)
    import rochpcg_amd as hip

    def log(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    G = hip.Problem(*dims, nlevels=4, stores=hip.STORE_F64 | hip.STORE_MIXED)
    log(f"setup {dims} on the GPU: {time.perf_counter() - t0:.2f} s")
    flops = G.flops_per_iteration()
    K, W = args.steps, args.warmup
    tuned = {}
    if args.shapes or not args.no_tune:
        from rochpcg_amd import autotune
        t0 = time.perf_counter()
        if args.shapes:
            table = autotune.load(args.shapes)
            if not autotune.apply(G, table):
                sys.exit(f"bench: no launch shapes for this problem in {args.shapes}")
            log(f"launch shapes from {args.shapes}")
        else:
            table = autotune.tune(G, mixed=False, log=log)
            table = autotune.tune(G, mixed=True, table=table, log=log)
            log(f"autotune: {time.perf_counter() - t0:.1f} s")
        if args.save_shapes:
            autotune.save(table, args.save_shapes)
        L0 = G.levels[0]
        tuned = {k: {op: v["shape"] for op, v in e.items()} for k, e in table.items()
                 if k.startswith(f"{L0.nx}x{L0.ny}x{L0.nz}:")}
    use_graph = not args.eager

    def timed(mode):
        G.cg_begin(mode=mode, max_iter=W + K, tol=0.0)
        G.cg_iterate(W, use_graph=use_graph)
        G.synchronize()
        t = time.perf_counter()
        G.cg_iterate(K, use_graph=use_graph)
        G.synchronize()
        el = time.perf_counter() - t
        res = G.cg_end()
        assert res["niters"] == W + K, res
        return el

    import numpy as np
    G.vec["x"].upload(np.zeros(G.m))
    el64 = timed(hip.CG_F64)
    log(f"FP64: {K} iterations in {el64 * 1e3:.1f} ms")
    G.vec["x"].upload(np.zeros(G.m))
    elmx = timed(hip.CG_MIXED)
    log(f"mixed: {K} iterations in {elmx * 1e3:.1f} ms")
    ms64 = el64 * 1e3 / K
    msmx = elmx * 1e3 / K
    # algorithmic bytes of this library's kernels (each ELL slot a kernel reads
    # once, neighbour gathers from cache, own-row vectors once)
    mg64, b64 = G.iteration_bytes(hip.CG_F64)
    _, bmx = G.iteration_bytes(hip.CG_MIXED)

    # per-phase HIP-event timing of K FP64 iterations (eager launches)
    G.vec["x"].upload(np.zeros(G.m))
    G.cg_begin(mode=hip.CG_F64, max_iter=K, tol=0.0)
    ph = G.cg_iterate_profiled(K)
    G.cg_end()
    spmv_ms = ph["spmv"] / K
    spmv_gbps = spmv_bytes(G.m) / (spmv_ms * 1e-3) / 1e9
    mg_ms = ph["mg"] / K
    mg_gbps = mg64 / (mg_ms * 1e-3) / 1e9
    log(f"phases (ms per iteration): " + ", ".join(f"{k} {v / K:.3f}" for k, v in ph.items()))

    out = {
        "metric": METRIC,
        "value": round(flops / (ms64 * 1e-3) / 1e9, 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(ms64, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"HPCG 27-point {dims[0]}x{dims[1]}x{dims[2]}, 4-level MG "
                               f"(1 pre/1 post multicolour SymGS), FP64 CG, b=A*ones, x0=0",
                   "nx": dims[0], "ny": dims[1], "nz": dims[2], "levels": 4,
                   "graph": use_graph, "parallelism": "single GPU", "level0_shapes": tuned},
        "hbm_TBps": round(b64 / (ms64 * 1e-3) / 1e12, 3),
        "bytes_per_iteration": int(b64),
        "mixed": {"value": round(flops / (msmx * 1e-3) / 1e9, 2), "unit": "GFLOP/s",
                  "ms_per_step": round(msmx, 4), "hbm_TBps": round(bmx / (msmx * 1e-3) / 1e12, 3),
                  "bytes_per_iteration": int(bmx), "speedup_per_iteration": round(ms64 / msmx, 3),
                  "dtype": "f64 outer CG / f32 vectors + f16 ELL in MG"},
        "phase_ms_per_iteration": {k: round(v / K, 4) for k, v in ph.items()},
        "roofline": {"kernel": "k_spmv (outer SpMV + fused p'Ap), level 0",
                     "bound": "hbm", "achieved": round(spmv_gbps, 1), "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": round(spmv_gbps / HBM_PEAK_GBPS, 4),
                     "traffic": profiled_traffic(dims), "bytes_per_launch": spmv_bytes(G.m),
                     "avg_launch_ms": round(spmv_ms, 4)},
        "mg_roofline": {"kernels": "FP64 V-cycle, all levels (sweeps, restriction, prolongation)",
                        "bound": "hbm", "achieved": round(mg_gbps, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                        "frac": round(mg_gbps / HBM_PEAK_GBPS, 4), "bytes_per_vcycle": int(mg64),
                        "ms_per_vcycle": round(mg_ms, 4)},
    }

    if not args.no_solve:
        # solve to 1e-9 from x0 = 0 with x* uniform in [-1,1] (splitmix64, --seed)
        G.make_rhs(1, args.seed)
        xs = G.vec["xexact"].host()
        sol = {}
        # "mixed_refine10": the mixed solve with the f64 true residual b - A x
        # recomputed every 10 iterations (iterative refinement); off in "mixed"
        for name, mode, every in (("f64", hip.CG_F64, 0), ("mixed", hip.CG_MIXED, 0),
                                  ("mixed_refine10", hip.CG_MIXED, 10)):
            G.set_residual_replacement(every)
            G.vec["x"].upload(np.zeros(G.m))
            G.synchronize()
            t = time.perf_counter()
            r = G.cg(max_iter=500, tol=1e-9, mode=mode, use_graph=use_graph, history=False)
            el = time.perf_counter() - t
            x = G.vec["x"].host()
            # true residual in f64 on the GPU: r = b - A x
            d = hip.DeviceArray(G.m)
            G.spmv(G.vec["x"], d)
            G.waxpby(1.0, G.vec["b"], -1.0, d, d)
            rt = np.sqrt(G.dot(d, d)) / np.sqrt(G.dot(G.vec["b"], G.vec["b"]))
            sol[name] = {"iterations": r["niters"], "seconds": round(el, 4),
                         "rel_residual": float(f"{r['normr'] / r['normr0']:.3e}"),
                         "true_rel_residual": float(f"{rt:.3e}"),
                         "err_vs_xstar": float(f"{np.linalg.norm(x - xs) / np.linalg.norm(xs):.3e}"),
                         "gflops": round(flops * r["niters"] / el / 1e9, 2)}
            log(f"solve {name}: {sol[name]}")
        G.set_residual_replacement(0)
        sol["speedup_time_to_solution"] = round(sol["f64"]["seconds"] / sol["mixed"]["seconds"], 3)
        sol["iteration_delta"] = sol["mixed"]["iterations"] - sol["f64"]["iterations"]
        out["solve_1e-9"] = sol
        sol["seed"] = args.seed
    G.close()

    if not args.no_batch:
        out["batched"] = batched_leg(hip, args.batch_count, log)

    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        log(f"CPU baseline on {threads} threads ...")
        out["cpu_baseline"] = cpu_baseline(dims, args.cpu_iters, threads)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
