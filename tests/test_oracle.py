"""Pins of the CPU oracle against closed forms, brute force and invariants.

These tests check the oracle (oracle/hpcg_oracle.c) itself -- the thing every
GPU parity test trusts -- against what the mathematics fixes: the 27-point
stencil's closed forms, dense brute force on tiny grids, the colour table that
the JPL rules and glibc's rand() draws imply, stable-sort semantics, the SymGS
triangular-solve identity and the symmetry of the MG preconditioner.
"""
import ctypes
import itertools

import numpy as np
import pytest

DIMS_SMALL = [(2, 2, 2), (3, 4, 5), (4, 4, 4), (8, 6, 4), (5, 3, 7)]


def dense_stencil(nx, ny, nz):
    """Brute-force 27-point operator: 26 on the diagonal, -1 for each neighbour."""
    m = nx * ny * nz
    A = np.zeros((m, m))
    for iz, iy, ix in itertools.product(range(nz), range(ny), range(nx)):
        i = (iz * ny + iy) * nx + ix
        for jz, jy, jx in itertools.product(range(nz), range(ny), range(nx)):
            if max(abs(jz - iz), abs(jy - iy), abs(jx - ix)) <= 1:
                j = (jz * ny + jy) * nx + jx
                A[i, j] = 26.0 if i == j else -1.0
    return A


def to_dense(cols, vals, nnz):
    m = len(nnz)
    A = np.zeros((m, m))
    for i in range(m):
        for k in range(nnz[i]):
            A[i, cols[i, k]] += vals[i, k]
    return A


@pytest.mark.parametrize("dims", DIMS_SMALL + [(16, 16, 16)])
def test_generate_closed_forms(oracle_mod, dims):
    nx, ny, nz = dims
    g = oracle_mod.generate(nx, ny, nz)
    iz, iy, ix = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    f = lambda i, n: 1 + (i > 0).astype(int) + (i < n - 1).astype(int)
    nnz_cf = (f(ix, nx) * f(iy, ny) * f(iz, nz)).ravel()
    assert np.array_equal(g["nnz"], nnz_cf)
    assert int(g["nnz"].sum()) == (3 * nx - 2) * (3 * ny - 2) * (3 * nz - 2)
    assert np.array_equal(g["b"], 27.0 - nnz_cf)
    assert np.all(g["xexact"] == 1.0)
    m = nx * ny * nz
    rows = np.arange(m)
    assert np.all(g["cols"][rows, g["diag"]] == rows)
    assert np.all(g["vals"][rows, g["diag"]] == 26.0)
    # natural stencil order is ascending column order; padding sits at the tail
    for i in range(0, m, max(1, m // 97)):
        c = g["cols"][i, :g["nnz"][i]]
        assert np.all(np.diff(c) > 0)
        assert np.all(g["cols"][i, g["nnz"][i]:] == -1)
        assert np.all(g["vals"][i, g["nnz"][i]:] == 0.0)
        assert np.sum(g["vals"][i, :g["nnz"][i]] == -1.0) == g["nnz"][i] - 1


@pytest.mark.parametrize("dims", [(2, 2, 2), (3, 2, 4), (4, 4, 4), (3, 5, 2)])
def test_generate_matches_dense_bruteforce(oracle_mod, dims):
    g = oracle_mod.generate(*dims)
    A = to_dense(g["cols"], g["vals"], g["nnz"])
    D = dense_stencil(*dims)
    assert np.array_equal(A, D)
    assert np.array_equal(A, A.T)
    assert np.array_equal(A @ np.ones(len(A)), g["b"])


@pytest.mark.parametrize("dims", [(4, 4, 4), (8, 6, 2), (16, 8, 4)])
def test_f2c_injection(oracle_mod, dims):
    nx, ny, nz = dims
    f2c = oracle_mod.f2c(nx, ny, nz)
    nxc, nyc, nzc = nx // 2, ny // 2, nz // 2
    k = np.arange(nxc * nyc * nzc)
    kx, ky, kz = k % nxc, (k // nxc) % nyc, k // (nxc * nyc)
    assert np.array_equal(f2c, (2 * kz) * nx * ny + (2 * ky) * nx + 2 * kx)
    # bijection onto the all-even fine points
    fx, fy, fz = f2c % nx, (f2c // nx) % ny, f2c // (nx * ny)
    assert np.all((fx % 2 == 0) & (fy % 2 == 0) & (fz % 2 == 0))
    assert len(np.unique(f2c)) == len(f2c) == (nx // 2) * (ny // 2) * (nz // 2)


@pytest.mark.parametrize("dims", [(4, 4, 4), (6, 4, 8), (5, 3, 7)])
def test_row_hash_orders_neighbours_by_class(oracle_mod, dims):
    nx, ny, nz = dims
    h = oracle_mod.row_hash(nx, ny, nz)
    g = oracle_mod.generate(nx, ny, nz)
    m = len(h)
    i = np.arange(m)
    ix, iy, iz = i % nx, (i // nx) % ny, i // (nx * ny)
    cls = ((ix & 1) << 2) | ((iy & 1) << 1) | (iz & 1)
    for r in range(m):
        for c in g["cols"][r, :g["nnz"][r]]:
            if c != r:
                assert cls[c] != cls[r]
                assert (h[c] > h[r]) == (cls[c] > cls[r])


def test_glibc_draws_of_the_reference_seed():
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(0x586744)
    draws = [libc.rand() % 8 for _ in range(8)]
    assert draws == [0, 7, 1, 2, 4, 3, 5, 6]


@pytest.mark.parametrize("dims", [(2, 2, 2), (4, 4, 4), (8, 6, 4), (5, 3, 7), (3, 3, 3), (16, 16, 16)])
def test_jpl_simulation_equals_closed_form(oracle_mod, dims):
    nx, ny, nz = dims
    g = oracle_mod.generate(nx, ny, nz)
    colors, ncol, rounds = oracle_mod.jpl(g["cols"], oracle_mod.row_hash(nx, ny, nz))
    assert ncol == 8
    assert np.array_equal(colors, oracle_mod.closed_form_colors(nx, ny, nz))
    # a proper colouring
    m = nx * ny * nz
    for r in range(0, m, max(1, m // 211)):
        nb = g["cols"][r, :g["nnz"][r]]
        nb = nb[nb != r]
        assert np.all(colors[nb] != colors[r])
    # round k of the reference records the count of the k-th drawn colour value
    drawn = [0, 7, 1, 2, 4, 3, 5, 6]
    for k, c in enumerate(drawn):
        assert rounds[k] == np.sum(colors == c)


@pytest.mark.parametrize("dims", [(4, 3, 2), (3, 4, 2)])
def test_jpl_equal_hash_levels_terminate(oracle_mod, dims):
    """4x3x2: neighbouring rows share a hash (|dcrd| = nx*ny+2nx+4 = N), where
    the reference's >=/<= rule never colours them; the tie-break by row index
    still yields a proper colouring."""
    nx, ny, nz = dims
    g = oracle_mod.generate(nx, ny, nz)
    colors, ncol, _ = oracle_mod.jpl(g["cols"], oracle_mod.row_hash(nx, ny, nz))
    assert ncol > 0 and np.all(colors >= 0)
    for r in range(nx * ny * nz):
        nb = g["cols"][r, :g["nnz"][r]]
        nb = nb[nb != r]
        assert np.all(colors[nb] != colors[r])


def test_perm_is_stable_sort(oracle_mod):
    rng = np.random.default_rng(3)
    colors = rng.integers(0, 8, 5000).astype(np.int32)
    perm, sizes, offsets = oracle_mod.perm_from_colors(colors, 8)
    order = np.argsort(colors, kind="stable")
    expect = np.empty_like(perm)
    expect[order] = np.arange(len(colors))
    assert np.array_equal(perm, expect)
    assert np.array_equal(sizes, np.bincount(colors, minlength=8))
    assert np.array_equal(offsets, np.concatenate([[0], np.cumsum(sizes)]))


@pytest.mark.parametrize("dims", [(4, 4, 4), (6, 4, 2)])
def test_permuted_matrix_is_PAPt(oracle_mod, dims):
    g = oracle_mod.generate(*dims)
    colors = oracle_mod.closed_form_colors(*dims)
    perm, _, _ = oracle_mod.perm_from_colors(colors, 8)
    pc, pv, pn, pd = oracle_mod.permute_matrix(g["cols"], g["vals"], g["nnz"], perm)
    A = to_dense(g["cols"], g["vals"], g["nnz"])
    Pm = np.zeros_like(A)
    Pm[perm, np.arange(len(perm))] = 1.0
    assert np.array_equal(to_dense(pc, pv, pn), Pm @ A @ Pm.T)
    for i in range(len(pn)):
        assert np.all(np.diff(pc[i, :pn[i]]) > 0)
        assert pc[i, pd[i]] == i


@pytest.mark.parametrize("order", [0, 1])
@pytest.mark.parametrize("dims", [(4, 4, 4), (8, 6, 4)])
def test_spmv_matches_dense(oracle_mod, order, dims):
    P = oracle_mod.Problem(*dims, nlevels=2, order=order)
    e = P.export(0)
    A = to_dense(e["cols"], e["vals"], e["nnz"])
    x = np.random.default_rng(1).uniform(-1, 1, P.m)
    y = P.spmv(x)
    assert np.allclose(y, A @ x, rtol=0, atol=1e-13)
    assert np.array_equal(P.spmv(np.ones(P.m)), P.rhs_ones())


def _symgs_operator(A):
    """M = (D+L) D^-1 (D+U): one symmetric GS step from x=0 solves M x = r."""
    D = np.diag(np.diag(A))
    L = np.tril(A, -1)
    U = np.triu(A, 1)
    return (D + L) @ np.linalg.inv(D) @ (D + U)


@pytest.mark.parametrize("order", [0, 1])
def test_symgs_triangular_identity(oracle_mod, order):
    P = oracle_mod.Problem(4, 4, 4, nlevels=1, order=order)
    e = P.export(0)
    A = to_dense(e["cols"], e["vals"], e["nnz"])
    r = np.random.default_rng(2).uniform(-1, 1, P.m)
    x = P.symgs(r, np.zeros(P.m))
    M = _symgs_operator(A)
    assert np.linalg.norm(M @ x - r) <= 1e-13 * np.linalg.norm(r)
    # a planted mistake (dropping the backward sweep) is caught by the same check
    D = np.diag(A)
    xf = np.zeros(P.m)
    for i in range(P.m):
        xf[i] = (r[i] - A[i, :i] @ xf[:i]) / D[i]
    assert np.linalg.norm(M @ xf - r) > 1e-3 * np.linalg.norm(r)


@pytest.mark.parametrize("order", [0, 1])
def test_mg_preconditioner_symmetry(oracle_mod, order):
    """TestSymmetry.cpp:88-166 on the oracle: departures scaled by eps <= 1."""
    P = oracle_mod.Problem(16, 16, 16, nlevels=4, order=order)
    rng = np.random.default_rng(5)
    x = rng.uniform(1, 2, P.m)
    y = rng.uniform(1, 2, P.m)
    xn, yn = x @ x, y @ y
    anorm = 2 * 26.0
    scale = (xn * anorm * yn + yn * anorm * xn) * np.finfo(float).eps
    assert abs(x @ P.spmv(y) - y @ P.spmv(x)) / scale <= 1.0
    assert abs(x @ P.mg(y) - y @ P.mg(x)) / scale <= 1.0


def test_cg_reaches_reference_tolerance_within_check_niter(oracle_mod):
    """test_rochpcg.cpp:88-119 / main.cpp:334-432: 50 natural-order CG_ref
    iterations define refTolerance; the multicolour CG must reach it within
    check_niter (53 for 1 rank at 32^3)."""
    nat = oracle_mod.Problem(32, nlevels=4, order=0)
    ref = nat.cg(nat.rhs_ones(), max_iter=50)
    ref_tol = ref["normr"] / ref["normr0"]
    col = oracle_mod.Problem(32, nlevels=4, order=1)
    opt = col.cg(col.rhs_ones(), max_iter=500, tol=ref_tol)
    assert opt["normr"] / opt["normr0"] <= ref_tol
    assert opt["niters"] <= 53


def test_testcg_expected_iterations(oracle_mod):
    """TestCG.cpp:111-177: diagonal x1e6 ((i+2)x1e6 for i<9): <= 12 unpreconditioned, <= 2 preconditioned."""
    P = oracle_mod.Problem(16, nlevels=4, order=1)
    d0 = P.copy_diagonal()
    b = P.rhs_ones()
    scale = np.full(P.m, 1.0e6)
    scale[:9] = (np.arange(9) + 2) * 1.0e6
    P.replace_diagonal(d0 * scale)
    bs = b * scale
    r0 = P.cg(bs, max_iter=50, tol=1e-12, precond=False)
    r1 = P.cg(bs, max_iter=50, tol=1e-12, precond=True)
    P.replace_diagonal(d0)
    assert r0["niters"] <= 12
    assert r1["niters"] <= 2


def test_testnorms_variance_rule():
    """TestNorms.cpp:31-46: pass iff the variance of the residual norms < 1e-6."""
    from rochpcg_amd.vv import testnorms
    assert testnorms([1e-9, 1e-9, 1.1e-9])["pass"]
    assert not testnorms([1.0, 0.0, 1.0, 0.0])["pass"]


def test_uniform_rng_is_splitmix64(oracle_mod):
    def splitmix(seed, i):
        mask = (1 << 64) - 1
        z = (seed + (i + 1) * 0x9E3779B97F4A7C15) & mask
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
        z ^= z >> 31
        return (z >> 11) * 2.0 ** -53
    for seed in (1234, 99, 7):
        for i in (0, 1, 12345, 10 ** 9):
            assert oracle_mod.uniform(seed, i) == splitmix(seed, i)


def test_rhs_random_is_A_times_xstar(oracle_mod):
    P = oracle_mod.Problem(8, 6, 4, nlevels=1, order=1)
    xs, b = P.rhs_random(1234)
    assert np.all(np.abs(xs) <= 1.0)
    assert np.allclose(P.spmv(xs), b, rtol=0, atol=1e-13)


def _golden(name):
    import os
    return np.load(os.path.join(os.path.dirname(__file__), "golden", name), allow_pickle=False)


@pytest.mark.parametrize("order,name", [(0, "natural"), (1, "jpl")])
def test_oracle_matches_golden_cg_history(oracle_mod, order, name):
    """Fixture written by tests/golden/make_golden.py from the reference's own
    CG_ref (oracle/_ref) and checked equal to the oracle there."""
    g = _golden("cg_history_32.npz")
    P = oracle_mod.Problem(32, nlevels=4, order=order)
    res = P.cg(P.rhs_ones(), max_iter=50)
    assert np.array_equal(res["history"], g[f"history_{name}"])


def test_oracle_matches_golden_kernels(oracle_mod):
    g = _golden("kernels_16.npz")
    P = oracle_mod.Problem(16, nlevels=4, order=1)
    x, r = g["x"], g["r"]
    assert np.array_equal(P.spmv(x), g["Ax"])
    assert np.array_equal(P.symgs(r, np.zeros(P.m)), g["symgs_zero"])
    assert np.array_equal(P.symgs(r, x.copy()), g["symgs_x"])
    assert np.array_equal(P.mg(r), g["mg"])


@pytest.mark.parametrize("n", [80, 96, 104])
def test_oracle_matches_golden_reference_tolerance(oracle_mod, n):
    """The reference tolerances at 80^3, 96^3 and 104^3 (50 natural-order
    CG_ref iterations, written from the reference's own _ref code) are
    reproduced by the oracle to 1e-10 relative (OpenMP reduction order aside)."""
    g = _golden("ref_tolerance.npz")
    tol = float(g["reduction"][list(g["n"]).index(n)])
    P = oracle_mod.Problem(n, nlevels=4, order=oracle_mod.ORDER_NATURAL)
    res = P.cg(P.rhs_ones(), max_iter=50)
    assert abs(res["normr"] / res["normr0"] - tol) <= 1e-10 * tol


# ---------------------------------------------------------------------------
# numpy restatements used by the full-size GPU checks, and the compensated dot

def test_compensated_dot_is_correctly_summed(oracle_mod):
    """DOT_COMPENSATED sums the rounded products to within 2 ulp of their
    exactly rounded sum (math.fsum); the reference's reduction is further off
    on long vectors of same-signed products."""
    import math
    rng = np.random.default_rng(31)
    x = rng.uniform(0.5, 1.5, 3_000_000)
    y = rng.uniform(0.5, 1.5, 3_000_000)
    exact = math.fsum(x * y)
    oracle_mod.set_dot_mode(oracle_mod.DOT_COMPENSATED)
    try:
        c = oracle_mod.dot(x, y)
        cc = oracle_mod.dot(x, x)
    finally:
        oracle_mod.set_dot_mode(oracle_mod.DOT_REFERENCE)
    assert abs(c - exact) <= 2 * np.spacing(exact)
    assert abs(cc - math.fsum(x * x)) <= 2 * np.spacing(cc)
    r = oracle_mod.dot(x, y)
    assert abs(r - exact) <= 1e-10 * exact       # one thread: sequential sum


def test_stencil_apply_equals_spmv(oracle_mod):
    """The numpy 27-point operator against ComputeSPMV_ref's restatement
    (natural order), slab by slab: to rounding, and exactly on integer data
    (A * ones = b)."""
    for dims in [(16, 16, 16), (24, 16, 32), (8, 40, 12), (5, 3, 7)]:
        P = oracle_mod.Problem(*dims, nlevels=1, order=oracle_mod.ORDER_NATURAL)
        x = np.random.default_rng(3).uniform(-1, 1, P.m)
        y = P.spmv(x)
        ys = oracle_mod.stencil_apply(x, *dims, chunk=3)
        assert np.max(np.abs(y - ys)) <= 1e-14 * np.max(np.abs(y))
        assert np.array_equal(oracle_mod.stencil_apply(np.ones(P.m), *dims), P.rhs_ones())
        xi = np.random.default_rng(4).integers(-100, 100, P.m).astype(np.float64)
        assert np.array_equal(oracle_mod.stencil_apply(xi, *dims), P.spmv(xi))


def test_uniform_vector_is_orc_uniform(oracle_mod):
    u = oracle_mod.uniform_vector(99, 5000)
    assert all(u[i] == oracle_mod.uniform(99, i) for i in range(5000))


@pytest.mark.parametrize("dims", [(16, 16, 16), (24, 16, 32), (6, 10, 4)])
def test_split_ell_rows_layout(oracle_mod, dims):
    """split_ell_rows against the oracle's tail-padded multicolour ELL: lower
    entries and the diagonal from slot 0, upper entries ending at slot 26,
    -1 / 0.0 between."""
    P = oracle_mod.Problem(*dims, nlevels=1, order=oracle_mod.ORDER_JPL)
    e = P.export(0)
    perm = e["perm"]
    iperm = np.empty_like(perm)
    iperm[perm] = np.arange(len(perm), dtype=np.int32)
    cols, vals = oracle_mod.split_ell_rows(*dims, perm, iperm, np.arange(P.m))
    oc, ov, d = e["cols"], e["vals"], e["diag"]
    ne = (oc >= 0).sum(axis=1)
    for r in range(P.m):
        nu = ne[r] - d[r] - 1
        want_c = np.full(27, -1)
        want_v = np.zeros(27)
        want_c[:d[r] + 1] = oc[r, :d[r] + 1]
        want_v[:d[r] + 1] = ov[r, :d[r] + 1]
        if nu:
            want_c[27 - nu:] = oc[r, d[r] + 1:ne[r]]
            want_v[27 - nu:] = ov[r, d[r] + 1:ne[r]]
        assert np.array_equal(cols[r], want_c), r
        assert np.array_equal(vals[r], want_v), r
