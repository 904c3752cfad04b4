"""CheckProblem on the device (hpcgx_check_problem): every row of every level
of the optimized form against the statements of the reference's CheckProblem
(CheckProblem.cpp:78-179) -- exactly the in-range 27-point neighbours through
the permutation, in ascending order, 26 / -1, the closed-form nonzero counts --
plus the permutation, the colouring and the split-padded layout.  The full-size
grids are checked in tests/test_gpu_configs.py."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ERRORS = ("perm", "color", "ell", "value", "desc", "vectors")


def _nnz_closed_form(nx, ny, nz):
    return (3 * nx - 2) * (3 * ny - 2) * (3 * nz - 2)


def check_all_levels(G):
    """Asserts a clean check of every level (vectors on the finest); returns
    the per-level results."""
    out = []
    for lv, L in enumerate(G.levels):
        c = G.check_problem(lv, vectors=(lv == 0))
        assert all(c[k] == 0 for k in ERRORS), (lv, c)
        assert c["rows"] == L.m, (lv, c)
        assert c["nnz"] == L.nnz == _nnz_closed_form(L.nx, L.ny, L.nz), (lv, c)
        out.append(c)
    return out


@pytest.mark.parametrize("dims", [(16, 16, 16), (32, 24, 16), (24, 16, 32), (40, 8, 16), (8, 8, 8), (256, 16, 16)])
def test_check_problem_clean(hip, dims):
    print(f"[gpu] check_problem {dims}", flush=True)
    G = hip.Problem(*dims, nlevels=4, stores=hip.STORE_F64 | hip.STORE_MIXED)
    check_all_levels(G)
    G.close()


def test_check_problem_single_store(hip):
    """Levels built with one value store only are checked on that store."""
    for stores in (hip.STORE_F64, hip.STORE_MIXED):
        G = hip.Problem(16, 16, 16, nlevels=4, stores=stores)
        check_all_levels(G)
        G.close()


def test_check_problem_sees_changed_values_and_vectors(hip):
    """The check follows the data: a diagonal replaced through the API shows
    up as value errors on every row (ELL diagonal slot, diagonal array, 1/a_ii),
    and is clean again once restored; a seeded right-hand side differs from
    the default one in b and xexact."""
    G = hip.Problem(16, 16, 16, nlevels=2, stores=hip.STORE_F64)
    m = G.m
    d = hip.DeviceArray(m)
    G.copy_diagonal(d)
    assert np.all(d.host() == 26.0)
    d2 = hip.DeviceArray.from_host(np.full(m, 52.0))
    G.replace_diagonal(d2)
    c = G.check_problem(0, vectors=True)
    assert c["value"] == 3 * m, c
    assert all(c[k] == 0 for k in ERRORS if k != "value"), c
    G.replace_diagonal(d)
    assert G.check_problem(0, vectors=True)["value"] == 0
    G.make_rhs(1, 5)
    c = G.check_problem(0, vectors=True)
    assert c["vectors"] >= m and c["ell"] == 0 and c["value"] == 0, c
    G.make_rhs(0)
    assert G.check_problem(0, vectors=True)["vectors"] == 0
    d.free()
    d2.free()
    G.close()
