"""The reference-named C++ layer (include/hpcg) driven like the reference's
main.cpp, on the GPU; checked against the oracle and the C-ABI path."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "build", "drop_in_main")


def run(*args):
    print(f"[gpu] drop_in_main {' '.join(map(str, args))}", flush=True)
    out = subprocess.run([EXE] + [str(a) for a in args], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    vals = {}
    for line in out.stdout.splitlines():
        parts = line.split()
        if len(parts) == 2:
            try:
                vals[parts[0]] = float(parts[1])
            except ValueError:
                pass
    return vals, out.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("dims", [(32, 32, 32), (24, 16, 32)])
def test_drop_in_layer(hip, oracle_mod, dims):
    v, _ = run(*dims)
    m = dims[0] * dims[1] * dims[2]
    assert v["done"] == 1
    assert v["rows"] == m and v["nonzeros"] == np.prod([3 * d - 2 for d in dims])
    assert v["nblocks"] == 8 and v["ublocks"] == 6
    assert v["spmv_residual"] == 0.0            # integers in f64: A*ones == b exactly
    assert v["restriction_diff"] <= 1e-15
    assert v["symgs_zero_diff"] <= 1e-14
    assert v["prolongation_norm2"] == pytest.approx(v["xc_norm2"], rel=1e-14)
    assert v["depsym_mg"] <= 1.0
    assert v["fused_dot_vs_dot"] <= 1e-13
    assert v["cg50_niters"] == 50
    assert v["cg_tol_scaled"] <= 1e-10 and v["cg_tol_niters"] <= 100
    assert v["solution_error"] < 1e-8
    assert v["testcg_noprec_niters"] <= 12 and v["testcg_prec_niters"] <= 2
    # the same solver through the C ABI gives the same 50-iteration residual
    G = hip.Problem(*dims, nlevels=4, stores=hip.STORE_F64)
    res = G.cg(max_iter=50)
    G.close()
    assert v["cg50_scaled"] == pytest.approx(res["normr"] / res["normr0"], rel=1e-12)
    # and the oracle (CG_ref in the same colour order) agrees to rounding noise
    O = oracle_mod.Problem(*dims, nlevels=4, order=oracle_mod.ORDER_JPL)
    ref = O.cg(O.rhs_ones(), max_iter=50)
    assert abs(v["cg50_scaled"] - ref["normr"] / ref["normr0"]) <= 1e-15


@pytest.mark.gpu
def test_drop_in_verbose_log_format(hip):
    _, out = run(16, 16, 16, "verbose")
    lines = [ln for ln in out.splitlines() if ln.startswith("HIP Iteration = ")]
    assert len(lines) == 50
    assert out.count("HIP Initial Residual = ") == 1
