"""The reference's natural-order path on the device (hpcgx_*_ref) against the
oracle in natural order: ComputeSPMV_ref, ComputeSYMGS_ref (sequential
Gauss-Seidel, run as wavefronts) and ComputeMG_ref bit for bit on every
level; CG_ref's residual history and the reference tolerance of the
benchmark driver (tests/golden/ref_tolerance.npz, written by the reference's
own _ref code) to the rounding of the dot products."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.mark.parametrize("dims", [(16, 16, 16), (32, 32, 32), (24, 16, 32), (32, 24, 16), (40, 8, 16)])
def test_ref_operators_bit_exact(hip, oracle_mod, dims):
    print(f"[gpu] natural-order operators {dims}", flush=True)
    G = hip.Problem(*dims, nlevels=4, stores=hip.STORE_F64)
    O = oracle_mod.Problem(*dims, nlevels=4, order=oracle_mod.ORDER_NATURAL)
    rng = np.random.default_rng(21)
    for lv in range(4):
        m = G.levels[lv].m
        x = rng.uniform(-1, 1, m)
        dx, dy = hip.DeviceArray.from_host(x), hip.DeviceArray(m)
        G.spmv_ref(dx, dy, level=lv)
        assert np.array_equal(dy.host(), O.spmv(x, lv)), f"spmv level {lv}"
        r = rng.uniform(-1, 1, m)
        x0 = rng.uniform(-1, 1, m)
        dr, dxx = hip.DeviceArray.from_host(r), hip.DeviceArray.from_host(x0)
        G.symgs_ref(dr, dxx, level=lv)
        assert np.array_equal(dxx.host(), O.symgs(r, x0.copy(), lv)), f"symgs level {lv}"
        dz = hip.DeviceArray(m)
        G.mg_ref(dr, dz, level=lv)
        assert np.array_equal(dz.host(), O.mg(r, lv)), f"mg level {lv}"
    G.close()


def test_ref_cg_history_32(hip):
    """CG_ref at 32^3 (b = A*ones, x0 = 0, 50 iterations) against the golden
    history of the reference's own _ref code.  Only the dot products round
    differently; the residual falls to 1e-20 of its start, so past rounding
    noise only an absolute bound applies: 1e-12 relative, or 1e-15 of the
    initial residual."""
    g = np.load(os.path.join(GOLDEN, "cg_history_32.npz"))
    hr = g["history_natural"]
    G = hip.Problem(32, nlevels=4, stores=hip.STORE_F64)
    res = G.cg_ref(max_iter=50)
    G.close()
    h = res["history"]
    assert res["niters"] == 50
    assert abs(h[0] - hr[0]) <= 1e-15 * hr[0]
    rel = np.abs(h - hr) / hr
    absd = np.abs(h - hr) / hr[0]
    print(f"[gpu] CG_ref 32^3: max rel {rel.max():.2e}, max |dh|/h0 {absd.max():.2e}", flush=True)
    assert np.all((rel <= 1e-12) | (absd <= 1e-15)), (rel, absd)


@pytest.mark.parametrize("n", [80, 104, 128, 256, 288])
def test_ref_tolerance_matches_reference(hip, n):
    """The 50-iteration residual reduction of CG_ref in natural order, as the
    benchmark driver takes it (main.cpp:334-344), against the reference's own
    _ref code: 1e-10 relative."""
    g = np.load(os.path.join(GOLDEN, "ref_tolerance.npz"))
    ref = float(g["reduction"][list(g["n"]).index(n)])
    print(f"[gpu] CG_ref {n}^3, 50 iterations", flush=True)
    G = hip.Problem(n, nlevels=4, stores=hip.STORE_F64)
    res = G.cg_ref(max_iter=50)
    G.close()
    red = res["normr"] / res["normr0"]
    print(f"[gpu] {n}^3: reduction {red:.10e} (reference {ref:.10e}), {res['seconds']:.2f} s", flush=True)
    assert res["niters"] == 50
    assert abs(red - ref) <= 1e-10 * ref


def test_ref_cg_unpreconditioned_and_user_vectors(hip, oracle_mod):
    """CG_ref without preconditioning and with a caller's b and x (natural
    order) against the oracle; tolerance stops it early."""
    dims = (24, 16, 32)
    G = hip.Problem(*dims, nlevels=4, stores=hip.STORE_F64)
    O = oracle_mod.Problem(*dims, nlevels=4, order=oracle_mod.ORDER_NATURAL)
    xs, b = O.rhs_random(5)
    db, dx = hip.DeviceArray.from_host(b), hip.DeviceArray(G.m)
    dx.upload(np.zeros(G.m))
    res = G.cg_ref(b=db, x=dx, max_iter=200, tol=1e-6, precond=False)
    ref = O.cg(b, max_iter=200, tol=1e-6, precond=False)
    assert res["niters"] == ref["niters"]
    h, hr = res["history"], ref["history"]
    assert np.all((np.abs(h - hr) <= 1e-12 * hr) | (np.abs(h - hr) <= 1e-15 * hr[0]))
    x = dx.host()
    assert np.linalg.norm(x - ref["x"]) <= 1e-10 * np.linalg.norm(ref["x"])
    G.close()
