#!/usr/bin/env python3
"""HPCG solver benchmark on one MI355X.

A step is one preconditioned CG iteration (4-level MG V-cycle with one pre-
and one post-smoothing multicolour SymGS per level, outer SpMV, dots and
vector updates) on the 27-point problem at 512^3 (134M rows, the largest
configuration, about 70 GB of HBM), all arrays resident in HBM.

Prints one JSON line: the FP64 HPCG-formula GFLOP/s (value), the mixed-
precision CG's GFLOP/s (f64 true residual every 10 iterations) and both
solves' time to a relative residual of 1e-9 (x* from splitmix64 seed 99), the
modelled HBM TB/s, a roofline entry for the outer SpMV kernel timed with HIP
events on the solver stream, a second leg at 320^3 (seed 1234), the
64-problem batched leg, and the CPU baseline (the reference's own CG_ref,
OpenMP, on this host's cores, on a bounded sample).
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


def vcycle_compulsory_bytes(level_rows):
    """FP64 V-cycle bytes by the compulsory-traffic model of the problem
    analysis (full-row ELL passes, x reused perfectly): per row of each level
    above the coarsest, 364 B zero-guess SymGS + 43 B fused residual and
    restriction + 4 B prolongation + 667.5 B post-smoothing SymGS; per row of
    the coarsest level, 364 B."""
    *upper, coarsest = level_rows
    return 1078.5 * sum(upper) + 364.0 * coarsest


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
    # every launch of 5 eager iterations between HIP events on the batch's
    # stream: per launch kind the time and the algorithmic bytes
    for name, mode in (("f64", hip.CG_F64), ("mixed", hip.CG_MIXED)):
        for P in B.problems:
            P.vec["x"].upload(np.zeros(P.m))
        prof = B.profile(iters=5, mode=mode)
        it_ms, it_bytes = prof.pop("iteration")
        kinds = {k: {"ms": round(ms, 4), "GBps": round(by / (ms * 1e-3) / 1e9, 1) if ms > 0 else None}
                 for k, (ms, by) in prof.items() if by > 0}
        dom = max(prof, key=lambda k: prof[k][0])
        dms, dby = prof[dom]
        res[f"roofline_{name}"] = {
            "kernel": f"batched {dom} launches (all levels and colours, one launch per colour for all problems)",
            "bound": "hbm", "achieved": round(dby / (dms * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBPS,
            "unit": "GB/s", "frac": round(dby / (dms * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4),
            "bytes_per_iteration": int(dby), "ms_per_iteration": round(dms, 4),
            "iteration_ms": round(it_ms, 4), "iteration_GBps": round(it_bytes / (it_ms * 1e-3) / 1e9, 1),
            "kinds": kinds}
        log(f"batched {name} profile: {res[f'roofline_{name}']}")
    B.close()
    return res


def cpu_info():
    """Model name of the host CPU (first /proc/cpuinfo entry)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def solve_leg(hip, G, seed, use_graph, log, modes):
    """Solves to 1e-9 from x0 = 0 with x* uniform in [-1,1] (splitmix64 `seed`),
    b = A x*: time, iterations, recursive and true (f64, b - A x on the GPU)
    relative residuals and the error against x*.  `modes`: (name, CG mode,
    true-residual interval)."""
    import numpy as np
    G.make_rhs(1, seed)
    xs = G.vec["xexact"].host()
    flops = G.flops_per_iteration()
    sol = {}
    for name, mode, every in modes:
        G.set_residual_replacement(every)
        G.vec["x"].upload(np.zeros(G.m))
        G.synchronize()
        t = time.perf_counter()
        r = G.cg(max_iter=1000, tol=1e-9, mode=mode, use_graph=use_graph, history=False)
        el = time.perf_counter() - t
        x = G.vec["x"].host()
        d = hip.DeviceArray(G.m)
        G.spmv(G.vec["x"], d)
        G.waxpby(1.0, G.vec["b"], -1.0, d, d)
        rt = np.sqrt(G.dot(d, d)) / np.sqrt(G.dot(G.vec["b"], G.vec["b"]))
        d.free()
        sol[name] = {"iterations": r["niters"], "seconds": round(el, 4),
                     "rel_residual": float(f"{r['normr'] / r['normr0']:.3e}"),
                     "true_rel_residual": float(f"{rt:.3e}"),
                     "err_vs_xstar": float(f"{np.linalg.norm(x - xs) / np.linalg.norm(xs):.3e}"),
                     "gflops": round(flops * r["niters"] / el / 1e9, 2)}
        log(f"solve {name}: {sol[name]}")
    G.set_residual_replacement(0)
    sol["seed"] = seed
    return sol


def tune_or_load(G, args, log):
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
    return {k: {op: v["shape"] for op, v in e.items()} for k, e in table.items()
            if k.startswith(f"{L0.nx}x{L0.ny}x{L0.nz}:")}


def timed(G, mode, K, W, use_graph):
    """Seconds of K CG iterations after W warm-up iterations (stream-synchronised)."""
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--nx", type=int, default=512)
    ap.add_argument("--ny", type=int, default=None)
    ap.add_argument("--nz", type=int, default=None)
    ap.add_argument("--seed", type=int, default=99, help="splitmix64 seed of x* in the solve-to-1e-9 leg")
    ap.add_argument("--second-nx", type=int, default=320,
                    help="cube side of the second leg (0 = none): per-iteration times and solves to 1e-9")
    ap.add_argument("--second-seed", type=int, default=1234)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-solve", action="store_true", help="skip the solve-to-1e-9 comparison")
    ap.add_argument("--cpu-iters", type=int, default=5)
    ap.add_argument("--cpu-nx", type=int, default=320, help="cube side of the bounded CPU-baseline sample")
    ap.add_argument("--eager", action="store_true", help="launch kernels one by one instead of a hipGraph")
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch shapes (256 threads, 1 row)")
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
    import numpy as np
    import rochpcg_amd as hip

    def log(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    G = hip.Problem(*dims, nlevels=4, stores=hip.STORE_F64 | hip.STORE_MIXED)
    setup_s = time.perf_counter() - t0
    log(f"setup {dims} on the GPU: {setup_s:.2f} s")
    G.make_rhs(1, args.seed)
    flops = G.flops_per_iteration()
    K, W = args.steps, args.warmup
    tuned = tune_or_load(G, args, log) if (args.shapes or not args.no_tune) else {}
    use_graph = not args.eager

    G.vec["x"].upload(np.zeros(G.m))
    el64 = timed(G, hip.CG_F64, K, W, use_graph)
    log(f"FP64: {K} iterations in {el64 * 1e3:.1f} ms")
    G.vec["x"].upload(np.zeros(G.m))
    elmx = timed(G, hip.CG_MIXED, K, W, use_graph)
    log(f"mixed: {K} iterations in {elmx * 1e3:.1f} ms")
    G.vec["x"].upload(np.zeros(G.m))
    G.set_residual_replacement(10)
    elmr = timed(G, hip.CG_MIXED, K, W, use_graph)
    G.set_residual_replacement(0)
    log(f"mixed, true residual every 10: {K} iterations in {elmr * 1e3:.1f} ms")
    ms64, msmx, msmr = (e * 1e3 / K for e in (el64, elmx, elmr))
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
    # the same time against the full-row compulsory byte model (the sweeps here
    # read only the slot windows they need, so they move fewer bytes than it counts)
    mg_comp = vcycle_compulsory_bytes([L.m for L in G.levels])
    log(f"phases (ms per iteration): " + ", ".join(f"{k} {v / K:.3f}" for k, v in ph.items()))
    # the box's read ceiling for the matrix stream the SpMV reads (the level-0
    # column and value arrays, 16-byte loads, no gathers), in the same run, and
    # the mixed-mode outer SpMV (f16 values, f64 vectors) timed as the CG runs it
    me = hip.ell_padded_rows(G.m)
    ceil64 = 27 * 12 * me / (G.time_operator(hip.OP_READ_STREAM, 0, mixed=False, reps=10) * 1e-3) / 1e9
    ceil16 = 27 * 6 * me / (G.time_operator(hip.OP_READ_STREAM, 0, mixed=True, reps=10) * 1e-3) / 1e9
    spmv16_ms = G.time_operator(hip.OP_SPMV, 0, mixed=True, reps=K)
    spmv16_gbps = spmv_bytes(G.m, 2) / (spmv16_ms * 1e-3) / 1e9
    log(f"read ceiling of the matrix stream: {ceil64:.0f} GB/s (f64 values), {ceil16:.0f} GB/s (f16 values); "
        f"mixed SpMV {spmv16_ms:.3f} ms, {spmv16_gbps:.0f} GB/s")

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
        "config": {"workload": f"HPCG 27-point {dims[0]}x{dims[1]}x{dims[2]} ({G.m} rows), 4-level MG "
                               f"(1 pre/1 post multicolour SymGS), x* uniform in [-1,1] (splitmix64 seed "
                               f"{args.seed}), b=A*x*, x0=0; FP64 CG per iteration (value) and the mixed "
                               f"FP32/FP16 V-cycle with the f64 true residual every 10 iterations",
                   "nx": dims[0], "ny": dims[1], "nz": dims[2], "levels": 4, "seed": args.seed,
                   "graph": use_graph, "parallelism": "single GPU", "level0_shapes": tuned},
        "setup_s": round(setup_s, 3),
        "hbm_TBps": round(b64 / (ms64 * 1e-3) / 1e12, 3),
        "bytes_per_iteration": int(b64),
        "mixed": {"value": round(flops / (msmr * 1e-3) / 1e9, 2), "unit": "GFLOP/s",
                  "ms_per_step": round(msmr, 4), "true_residual_every": 10,
                  "ms_per_step_no_refinement": round(msmx, 4),
                  "hbm_TBps": round(bmx / (msmx * 1e-3) / 1e12, 3),
                  "bytes_per_iteration": int(bmx), "speedup_per_iteration": round(ms64 / msmr, 3),
                  "dtype": "f64 outer CG / f32 vectors + f16 ELL in MG"},
        "phase_ms_per_iteration": {k: round(v / K, 4) for k, v in ph.items()},
        "roofline": {"kernel": "k_spmv (outer SpMV + fused p'Ap), level 0",
                     "bound": "hbm", "achieved": round(spmv_gbps, 1), "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": round(spmv_gbps / HBM_PEAK_GBPS, 4),
                     "traffic": profiled_traffic(dims), "bytes_per_launch": spmv_bytes(G.m),
                     "avg_launch_ms": round(spmv_ms, 4),
                     "read_ceiling_GBps": round(ceil64, 1), "frac_of_read_ceiling": round(spmv_gbps / ceil64, 4)},
        "mixed_spmv_roofline": {"kernel": "k_spmv (f16 values, f64 vectors, fused p'Ap), level 0",
                                "bound": "hbm", "achieved": round(spmv16_gbps, 1), "peak": HBM_PEAK_GBPS,
                                "unit": "GB/s", "frac": round(spmv16_gbps / HBM_PEAK_GBPS, 4),
                                "bytes_per_launch": spmv_bytes(G.m, 2), "avg_launch_ms": round(spmv16_ms, 4),
                                "read_ceiling_GBps": round(ceil16, 1),
                                "frac_of_read_ceiling": round(spmv16_gbps / ceil16, 4)},
        "mg_roofline": {"kernels": "FP64 V-cycle, all levels (sweeps, restriction, prolongation)",
                        "bound": "hbm", "achieved": round(mg_gbps, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                        "frac": round(mg_gbps / HBM_PEAK_GBPS, 4), "bytes_per_vcycle": int(mg64),
                        "ms_per_vcycle": round(mg_ms, 4),
                        "compulsory_bytes_per_vcycle": int(mg_comp),
                        "frac_compulsory_model": round(mg_comp / (mg_ms * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4)},
    }

    modes = (("f64", hip.CG_F64, 0), ("mixed_refine10", hip.CG_MIXED, 10), ("mixed", hip.CG_MIXED, 0))
    if not args.no_solve:
        sol = solve_leg(hip, G, args.seed, use_graph, log, modes)
        # the configuration's mixed solve: the f64 true residual every 10 iterations
        sol["speedup_time_to_solution"] = round(sol["f64"]["seconds"] / sol["mixed_refine10"]["seconds"], 3)
        sol["speedup_time_to_solution_no_refinement"] = round(sol["f64"]["seconds"] / sol["mixed"]["seconds"], 3)
        sol["iteration_delta"] = sol["mixed_refine10"]["iterations"] - sol["f64"]["iterations"]
        out["solve_1e-9"] = sol
    G.close()

    if args.second_nx:
        n2 = args.second_nx
        t0 = time.perf_counter()
        G2 = hip.Problem(n2, n2, n2, nlevels=4, stores=hip.STORE_F64 | hip.STORE_MIXED)
        G2.make_rhs(1, args.second_seed)
        log(f"second leg {n2}^3: setup {time.perf_counter() - t0:.2f} s")
        if not args.no_tune:
            from rochpcg_amd import autotune
            table = autotune.tune(G2, mixed=False)
            autotune.tune(G2, mixed=True, table=table)
        f2 = G2.flops_per_iteration()
        leg = {"workload": f"HPCG 27-point {n2}^3, 4-level MG, x* seed {args.second_seed}"}
        for name, mode, every in (("f64", hip.CG_F64, 0), ("mixed_refine10", hip.CG_MIXED, 10)):
            G2.set_residual_replacement(every)
            G2.vec["x"].upload(np.zeros(G2.m))
            el = timed(G2, mode, K, W, use_graph)
            leg[name] = {"ms_per_step": round(el * 1e3 / K, 4), "gflops": round(f2 * K / el / 1e9, 2)}
        G2.set_residual_replacement(0)
        leg["speedup_per_iteration"] = round(leg["f64"]["ms_per_step"] / leg["mixed_refine10"]["ms_per_step"], 3)
        if not args.no_solve:
            sol2 = solve_leg(hip, G2, args.second_seed, use_graph, log, modes[:2])
            sol2["speedup_time_to_solution"] = round(sol2["f64"]["seconds"] / sol2["mixed_refine10"]["seconds"], 3)
            leg["solve_1e-9"] = sol2
        G2.close()
        out[f"leg_{n2}"] = leg
        log(f"second leg: {leg}")

    if not args.no_batch:
        out["batched"] = batched_leg(hip, args.batch_count, log)

    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        log(f"CPU baseline on {threads} threads ...")
        cdims = (args.cpu_nx,) * 3
        cb = cpu_baseline(cdims, args.cpu_iters, threads)
        cb["cpu_model"] = cpu_info()
        cb["omp_proc_bind"] = os.environ.get("OMP_PROC_BIND", "unset (OpenMP default)")
        cb["omp_num_threads"] = threads
        if cdims != dims:
            cb["sample"] += (f"; a bounded sample of the same operator: the {dims[0]}x{dims[1]}x{dims[2]} problem "
                             f"in the reference's host format needs about 100 GB and minutes of setup")
        out["cpu_baseline"] = cb
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
