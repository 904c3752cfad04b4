"""Verification & validation of a run, on the GPU problem (reference: TestCG.cpp,
TestSymmetry.cpp, TestNorms.cpp, test_rochpcg.cpp:88-119, ReportResults.cpp:65-94).

The drivers call only the HIP library through ``rochpcg_amd.Problem``; the
rules (thresholds, scalings, departure formulas) are the reference's.
"""
from __future__ import annotations

import numpy as np


def testnorms(values):
    """TestNorms.cpp:31-46: mean via deltas to the first sample, population
    variance, pass iff variance < 1e-6."""
    v = np.asarray(values, dtype=np.float64)
    n = len(v)
    mean = v[0] + np.sum(v - v[0]) / n
    var = float(np.sum((v - mean) ** 2) / n)
    return dict(mean=float(mean), variance=var, samples=n, **{"pass": var < 1.0e-6})


def check_niter(n: int, ranks: int = 1) -> int:
    """Optimized-CG iteration ceilings of test_rochpcg.cpp:88-119 for cubes of side n."""
    if ranks == 1:
        if n in (16, 32, 104):
            return 53
        if n == 256:
            return 52
        if n == 288:
            return 51
        return 50
    return 56


def hpcg_flops_per_iteration(dims, levels=4, pre=1, post=1):
    """Flops of one CG iteration by the HPCG model (ReportResults.cpp:73-89)."""
    d = list(dims)
    nrow = d[0] * d[1] * d[2]
    nnz = []
    for _ in range(levels):
        nnz.append((3 * d[0] - 2) * (3 * d[1] - 2) * (3 * d[2] - 2))
        d = [v // 2 for v in d]
    f = 6.0 * nrow + 6.0 * nrow + 2.0 * nnz[0]
    for lv in range(levels - 1):
        f += pre * 4.0 * nnz[lv] + 2.0 * nnz[lv] + post * 4.0 * nnz[lv]
    f += 4.0 * nnz[levels - 1]
    return f


def testsymmetry(prob, seed=0):
    """TestSymmetry.cpp:88-166 on the GPU operators: |x'Ay - y'Ax| and
    |x'M^-1 y - y'M^-1 x| scaled by (2*||x||^2*||A||*||y||^2)*eps; pass iff <= 1.
    x, y uniform in [1, 2) (HIPFillRandomVector, Vector.hpp:108-120)."""
    from . import DeviceArray
    rng = np.random.default_rng(seed)
    m = prob.m
    x = DeviceArray.from_host(rng.uniform(1.0, 2.0, m))
    y = DeviceArray.from_host(rng.uniform(1.0, 2.0, m))
    z = DeviceArray(m)
    anorm = 2 * 26.0
    yn = prob.dot(y, y)
    prob.spmv(y, z)
    xtay = prob.dot(x, z)
    xn = prob.dot(x, x)
    prob.spmv(x, z)
    ytax = prob.dot(y, z)
    eps = np.finfo(np.float64).eps
    scale = (xn * anorm * yn + yn * anorm * xn) * eps
    dep_spmv = abs(xtay - ytax) / scale
    prob.mg(y, z)
    xtmy = prob.dot(x, z)
    prob.mg(x, z)
    ytmx = prob.dot(y, z)
    dep_mg = abs(xtmy - ytmx) / scale
    fails = int(dep_spmv > 1.0) + int(dep_mg > 1.0)
    return dict(depsym_spmv=float(dep_spmv), depsym_mg=float(dep_mg), count_fail=fails)


def testcg(prob):
    """TestCG.cpp:111-177: exaggerate the diagonal and b (x1e6, (i+2)x1e6 for
    position i < 9), run 2 unpreconditioned and 2 preconditioned CG calls to
    1e-12; expect <= 12 and <= 2 iterations; restore the matrix and b."""
    from . import DeviceArray
    m = prob.m
    d0 = DeviceArray(m)
    dx = DeviceArray(m)
    b0 = DeviceArray(m)
    prob.copy_diagonal(d0)
    prob.synchronize()
    dx.upload(d0.host())
    b = prob.vec["b"]
    b0.upload(b.host())
    prob.testcg_scale(dx, b)
    prob.replace_diagonal(dx)
    out = dict(count_pass=0, count_fail=0, niters_max_no_prec=0, niters_max_prec=0,
               expected_niters_no_prec=12, expected_niters_prec=2)
    try:
        for k in range(2):
            expected = 12 if k == 0 else 2
            for _ in range(2):
                prob.vec["x"].upload(np.zeros(m))
                res = prob.cg(max_iter=50, tol=1.0e-12, precond=(k == 1), use_graph=False)
                if res["niters"] <= expected:
                    out["count_pass"] += 1
                else:
                    out["count_fail"] += 1
                key = "niters_max_prec" if k == 1 else "niters_max_no_prec"
                out[key] = max(out[key], res["niters"])
                out["normr"] = res["normr"]
    finally:
        prob.replace_diagonal(d0)
        b.upload(b0.host())
        prob.vec["x"].upload(np.zeros(m))
    return out


def splitmix64(seed, i):
    """Output i of the splitmix64 generator seeded with `seed` (as an int)."""
    mask = (1 << 64) - 1
    z = (seed + (i + 1) * 0x9E3779B97F4A7C15) & mask
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
    return z ^ (z >> 31)


def batched_dims(count=64, seed=7, lo=16, hi=96, step=8):
    """Grid of each problem of the batched workload: nx, ny, nz drawn
    independently from {lo, lo+step, ..., hi}; dimension d of problem k takes
    splitmix64(seed) output 3k+d modulo the number of choices."""
    choices = list(range(lo, hi + 1, step))
    return [tuple(choices[splitmix64(seed, 3 * k + d) % len(choices)] for d in range(3))
            for k in range(count)]
