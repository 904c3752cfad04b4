"""Writes the golden fixtures of tests/golden/ (run once, in a tree with the
reference sources, after `make -C oracle oracle ref`).

Every vector is computed twice -- by the oracle's restatement and by the
reference's own _ref sources built into oracle/_ref -- with one OpenMP thread,
and written only if both agree bit for bit.  The fixtures then pin the oracle
wherever oracle/_ref cannot be built.

  cg_history_32.npz   CG_ref, 32^3, 4 levels, b = A*ones, x0 = 0, 50 iterations:
                      normr per iteration in natural order and in the JPL
                      multicolour order.
  kernels_16.npz      16x16x16 (JPL order): x ~ U[-1,1] (numpy default_rng(2024)),
                      A x, one SymGS from x = 0 and from x, one V-cycle.
  cg_history_320.npz  CG_ref at 320^3 in the JPL multicolour order, b = A*ones,
                      50 iterations, 8 threads (`make_golden.py history-large`;
                      iterations 0..5 checked against the reference's own _ref
                      code to 1e-13), and the same run with compensated dot
                      products (`history-large-compensated`).
  ref_tolerance.npz   the reference tolerance of main.cpp:336-345 -- normr/normr0
                      after 50 natural-order CG_ref iterations, b = A*ones,
                      x0 = 0, 4 levels -- for the cube sides of the reference's
                      own test (test_rochpcg.cpp:65-77) whose reduction stays
                      above rounding noise: 80 ... 128, 256, 280, 288 (16 and
                      32 reduce below 1e-20, 48 to 3e-15 and 64 to 1e-11,
                      where runs on 8 threads already differ by 1e-10)
                      (written by `make_golden.py ref-tolerance`; 8 threads,
                      oracle and reference within 1e-10 relative, the
                      reference's value kept).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import oracle  # noqa: E402


def agree(a, b, what):
    if not np.array_equal(a, b):
        raise SystemExit(f"oracle and reference disagree on {what}")
    return a


def main():
    oracle.build(ref=True)
    oracle.set_threads(1)
    out = {}
    for name, order in (("natural", oracle.ORDER_NATURAL), ("jpl", oracle.ORDER_JPL)):
        P = oracle.Problem(32, nlevels=4, order=order)
        R = oracle.RefHierarchy(P)
        b = P.rhs_ones()
        h = P.cg(b, max_iter=50)["history"]
        out[f"history_{name}"] = agree(h, R.history(b, max_iter=50), f"history {name}")
    np.savez(os.path.join(HERE, "cg_history_32.npz"), **out)

    P = oracle.Problem(16, nlevels=4, order=oracle.ORDER_JPL)
    R = oracle.RefHierarchy(P)
    rng = np.random.default_rng(2024)
    x = rng.uniform(-1, 1, P.m)
    r = rng.uniform(-1, 1, P.m)
    k = dict(x=x, r=r)
    k["Ax"] = agree(P.spmv(x), R.spmv(x), "spmv")
    k["symgs_zero"] = agree(P.symgs(r, np.zeros(P.m)), R.symgs(r, np.zeros(P.m)), "symgs from 0")
    k["symgs_x"] = agree(P.symgs(r, x.copy()), R.symgs(r, x.copy()), "symgs from x")
    k["mg"] = agree(P.mg(r), R.mg(r), "mg")
    np.savez(os.path.join(HERE, "kernels_16.npz"), **k)
    print("golden fixtures written")


def ref_tolerance(sizes=(80, 96, 104, 112, 120, 128, 256, 280, 288)):
    oracle.build(ref=True)
    oracle.set_threads(8)
    out = {"n": np.array(sizes), "reduction": np.zeros(len(sizes))}
    for i, n in enumerate(sizes):
        P = oracle.Problem(n, nlevels=4, order=oracle.ORDER_NATURAL)
        b = P.rhs_ones()
        o = P.cg(b, max_iter=50)
        R = oracle.RefHierarchy(P)
        r = R.cg(b, max_iter=50)
        R.close()
        # 8 OpenMP threads: the two codes' dot reductions round differently,
        # so they must agree to 1e-10 relative here, not bit for bit
        ro, rr = o["normr"] / o["normr0"], r["normr"] / r["normr0"]
        if not abs(ro - rr) <= 1e-10 * rr:
            raise SystemExit(f"oracle and reference disagree on the reference tolerance at {n}: {ro} vs {rr}")
        out["reduction"][i] = rr
        print(f"n={n}: 50-iteration reduction {rr:.6e} (oracle {ro:.6e})", flush=True)
        del P
    np.savez(os.path.join(HERE, "ref_tolerance.npz"), **out)


def history_large_compensated(n=320, iters=50):
    """Adds history_jpl_compensated to cg_history_<n>.npz: the same CG_ref with
    every dot product summed by the oracle's compensated summation instead of
    the reference's OpenMP reduction (oracle.DOT_COMPENSATED).  The difference
    between the two histories is the reference summation's own rounding error
    at this vector length."""
    import time
    oracle.build(ref=False)
    oracle.set_threads(8)
    oracle.set_dot_mode(oracle.DOT_COMPENSATED)
    t0 = time.time()
    P = oracle.Problem(n, nlevels=4, order=oracle.ORDER_JPL)
    o = P.cg(P.rhs_ones(), max_iter=iters)
    oracle.set_dot_mode(oracle.DOT_REFERENCE)
    h = o["history"]
    assert o["niters"] == iters and len(h) == iters + 1
    path = os.path.join(HERE, f"cg_history_{n}.npz")
    old = dict(np.load(path))
    rel = np.abs(h - old["history_jpl"]) / old["history_jpl"]
    print(f"{iters} iterations {time.time() - t0:.0f} s; against the reference summation: max relative "
          f"difference {rel.max():.2e}", flush=True)
    old["history_jpl_compensated"] = h
    np.savez(path, **old)


def history_large(n=320, iters=50, check=5, threads=8):
    """cg_history_<n>.npz: CG_ref in the JPL multicolour order at n^3 (4 levels,
    b = A*ones, x0 = 0, `iters` iterations, 8 OpenMP threads): normr per
    iteration.  The first `check` iterations are recomputed by the reference's
    own _ref code over the same arrays and must agree to 1e-13 relative (the
    OpenMP dot reductions of the two builds may combine thread partials in a
    different order)."""
    import time
    oracle.build(ref=True)
    oracle.set_threads(threads)
    t0 = time.time()
    P = oracle.Problem(n, nlevels=4, order=oracle.ORDER_JPL)
    b = P.rhs_ones()
    print(f"{n}^3 oracle setup {time.time() - t0:.0f} s", flush=True)
    t0 = time.time()
    o = P.cg(b, max_iter=iters)
    h = o["history"]
    print(f"{iters} iterations {time.time() - t0:.0f} s; reduction {h[-1] / h[0]:.6e}", flush=True)
    assert o["niters"] == iters and len(h) == iters + 1
    R = oracle.RefHierarchy(P)
    hr = R.history(b, max_iter=check)
    R.close()
    rel = np.abs(hr - h[:check + 1]) / h[:check + 1]
    print(f"reference _ref, iterations 0..{check}: max relative difference {rel.max():.2e}", flush=True)
    if not np.all(rel <= 1e-13):
        raise SystemExit(f"oracle and reference disagree on the {n}^3 history: {rel}")
    np.savez(os.path.join(HERE, f"cg_history_{n}.npz"), n=np.array([n, n, n]), history_jpl=h,
             threads=np.array(threads), ref_checked=np.array(check))


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "ref-tolerance":
        ref_tolerance()
    elif len(sys.argv) > 1 and sys.argv[1] == "history-large":
        history_large(int(sys.argv[2]) if len(sys.argv) > 2 else 320)
    elif len(sys.argv) > 1 and sys.argv[1] == "history-large-compensated":
        history_large_compensated(int(sys.argv[2]) if len(sys.argv) > 2 else 320)
    else:
        main()
