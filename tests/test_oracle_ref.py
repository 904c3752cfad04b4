"""The oracle against the reference's own code, bit for bit.

oracle/_ref/libhpcg_ref.so is built (make -C oracle ref) from ROCm/rocHPCG's
unmodified host sources -- GenerateProblem_ref, SetupHalo_ref, ComputeSPMV_ref,
ComputeSYMGS_ref, ComputeMG_ref, ComputeRestriction_ref,
ComputeProlongation_ref, ComputeDotProduct_ref, ComputeWAXPBY_ref, CG_ref --
with the same floating-point flags as the oracle.  Equal bits here mean the
restatement follows the reference's loops and expression order exactly.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.skipif(not oracle.ref_available(),
                                reason="oracle/_ref not built (needs the reference tree)")


@pytest.mark.parametrize("dims", [(2, 2, 2), (8, 6, 4), (16, 16, 16), (7, 5, 3)])
def test_generate_problem_ref_bits(oracle_mod, dims):
    g = oracle_mod.generate(*dims)
    r = oracle_mod.ref_generate(*dims)
    for k in ("cols", "vals", "nnz", "diag", "b", "xexact"):
        assert np.array_equal(g[k], r[k]), k
    assert np.all(r["x"] == 0.0)


@pytest.mark.parametrize("order", [0, 1])
@pytest.mark.parametrize("dims", [(16, 16, 16), (16, 8, 24)])
def test_kernels_bit_exact(oracle_mod, order, dims):
    P = oracle_mod.Problem(*dims, nlevels=4, order=order)
    R = oracle_mod.RefHierarchy(P)
    rng = np.random.default_rng(11)
    for lv in range(P.nlevels):
        m = P.levels[lv].m
        x = rng.uniform(-1, 1, m)
        assert np.array_equal(P.spmv(x, lv), R.spmv(x, lv))
        r = rng.uniform(-1, 1, m)
        x0 = rng.uniform(-1, 1, m)
        assert np.array_equal(P.symgs(r, x0.copy(), lv), R.symgs(r, x0.copy(), lv))
    r = rng.uniform(-1, 1, P.m)
    assert np.array_equal(P.mg(r), R.mg(r))
    a, b = rng.uniform(-1, 1, P.m), rng.uniform(-1, 1, P.m)
    assert oracle_mod.dot(a, b) == oracle.reflib().ref_dot(P.m, a.ctypes.data, b.ctypes.data)
    for alpha, beta in ((1.0, -0.5), (0.25, 1.0), (0.3, 0.7)):
        w = np.empty(P.m)
        oracle.reflib().ref_waxpby(P.m, alpha, a, beta, b, w)
        assert np.array_equal(oracle_mod.waxpby(alpha, a, beta, b), w)


@pytest.mark.parametrize("order", [0, 1])
def test_cg_history_bit_exact(oracle_mod, order):
    P = oracle_mod.Problem(16, nlevels=4, order=order)
    R = oracle_mod.RefHierarchy(P)
    b = P.rhs_ones()
    ours = P.cg(b, max_iter=20)
    hist = R.history(b, max_iter=20)
    assert np.array_equal(ours["history"], hist)
    ref = R.cg(b, max_iter=20)
    assert np.array_equal(ours["x"], ref["x"])
    # unpreconditioned path too
    ours = P.cg(b, max_iter=10, precond=False)
    ref = R.cg(b, max_iter=10, precond=False)
    assert ours["normr"] == ref["normr"] and np.array_equal(ours["x"], ref["x"])
