"""Batched uneven solves: every problem of a batch converges exactly as it
does alone (same iteration count, residual history equal up to the grouping
of the dot partials), in FP64 and in mixed precision, graph or eager."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DIMS = [(16, 16, 16), (24, 16, 32), (32, 24, 16), (16, 32, 24), (40, 16, 16), (16, 16, 48)]
# coarsest levels of 38, 8 and 22 colours: colour launches that only some
# problems of the batch take part in
MANY_COLOURS = [(256, 16, 16), (16, 16, 16), (128, 16, 16)]


def _rhs_seed(k):
    return 7 + k


@pytest.mark.parametrize("dims_list", [DIMS, MANY_COLOURS], ids=["uneven", "many_colours"])
@pytest.mark.parametrize("mode", [0, 1])
def test_batched_equals_individual(hip, mode, dims_list):
    print(f"[gpu] batch of {len(dims_list)} problems, mode {mode}", flush=True)
    B = hip.Batch(dims_list, nlevels=4, stores=hip.STORE_F64 | hip.STORE_MIXED)
    for k, P in enumerate(B.problems):
        P.make_rhs(1, _rhs_seed(k))
    res = B.cg(max_iter=300, tol=1e-9, mode=mode)
    for k, dims in enumerate(dims_list):
        P = B.problems[k]
        x = P.vec["x"].host()
        G = hip.Problem(*dims, nlevels=4, stores=hip.STORE_F64 | hip.STORE_MIXED)
        G.make_rhs(1, _rhs_seed(k))
        single = G.cg(max_iter=300, tol=1e-9, mode=mode)
        xs = G.vec["x"].host()
        xstar = G.vec["xexact"].host()
        G.close()
        assert res["normr"][k] / res["normr0"][k] <= 1e-9
        assert abs(int(res["niters"][k]) - single["niters"]) <= 1, (k, res["niters"][k], single["niters"])
        assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-8
        assert np.linalg.norm(x - xstar) / np.linalg.norm(xstar) < 1e-6
        # true residual of the batched solution
        d = hip.DeviceArray(P.m)
        P.spmv(P.vec["x"], d)
        P.waxpby(1.0, P.vec["b"], -1.0, d, d)
        assert np.sqrt(P.dot(d, d) / P.dot(P.vec["b"], P.vec["b"])) <= 2e-9
    B.close()


def test_batched_graph_equals_eager(hip):
    B = hip.Batch(DIMS[:4], nlevels=4, stores=hip.STORE_F64)
    for k, P in enumerate(B.problems):
        P.make_rhs(1, _rhs_seed(k))
    a = B.cg(max_iter=25, use_graph=False)
    xa = [P.vec["x"].host() for P in B.problems]
    for P in B.problems:
        P.vec["x"].upload(np.zeros(P.m))
    b = B.cg(max_iter=25, use_graph=True)
    assert np.array_equal(a["normr"], b["normr"])
    for P, xv in zip(B.problems, xa):
        assert np.array_equal(P.vec["x"].host(), xv)
    # block sizes change the launch shape, not the per-row arithmetic
    B.set_block(0, 128)
    for P in B.problems:
        P.vec["x"].upload(np.zeros(P.m))
    c = B.cg(max_iter=25, use_graph=True)
    # equal up to the grouping of dot partials; the 16^3 problem is already at
    # rounding noise (1e-16 relative), where only an absolute bound is meaningful
    diff = np.abs(c["normr"] - a["normr"])
    assert np.all((diff <= 1e-10 * a["normr"]) | (diff <= 1e-15 * a["normr0"]))
    xc = [P.vec["x"].host() for P in B.problems]
    # two rows per thread in the sweeps (where every segment allows it) leave
    # the per-row arithmetic and the reductions as they were: bit for bit
    for lv in range(4):
        B.set_rows(lv, 2)
    for P in B.problems:
        P.vec["x"].upload(np.zeros(P.m))
    d = B.cg(max_iter=25, use_graph=True)
    assert np.array_equal(d["normr"], c["normr"])
    for P, xv in zip(B.problems, xc):
        assert np.array_equal(P.vec["x"].host(), xv)
    B.close()


def test_batched_profile(hip):
    """Per-launch HIP-event timing of the batched iteration: every kind that
    runs has a time and a byte count, the kinds' times add up to no more than
    the iteration, and the SpMV's bytes are 27 x 12 + 16 per row (f64)."""
    B = hip.Batch(DIMS[:4], nlevels=4, stores=hip.STORE_F64 | hip.STORE_MIXED)
    rows = sum(a * b * c for a, b, c in DIMS[:4])
    for mode in (hip.CG_F64, hip.CG_MIXED):
        for P in B.problems:
            P.vec["x"].upload(np.zeros(P.m))
        prof = B.profile(iters=3, mode=mode)
        it_ms, it_bytes = prof.pop("iteration")
        assert it_ms > 0 and it_bytes > 0
        assert sum(ms for ms, _ in prof.values()) <= it_ms * 1.001
        for k in ("spmv", "fwd_zero", "bwd_zero", "fwd", "bwd", "restrict", "prolong", "dot", "cg_p", "cg_xr"):
            assert prof[k][0] > 0 and prof[k][1] > 0, k
        spmv_row = 27 * (12 if mode == hip.CG_F64 else 6) + 16
        assert prof["spmv"][1] == rows * spmv_row
    # the solver still runs normally afterwards
    for P in B.problems:
        P.vec["x"].upload(np.zeros(P.m))
    res = B.cg(max_iter=300, tol=1e-9)
    assert np.all(res["normr"] / res["normr0"] <= 1e-9)
    B.close()
