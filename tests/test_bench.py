"""The benchmark script's contract on the CPU: one GPU only, the byte models it
reports, and the committed traffic summaries it reads."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    out = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 2
    line = json.loads(out.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == 2 and "one GPU" in line["error"]


def test_spmv_bytes_per_row():
    # 27 slots of (int32 column + value), x and y once per row
    assert bench.spmv_bytes(1) == 27 * 12 + 16 == 340
    assert bench.spmv_bytes(1, vb=2) == 27 * 6 + 16 == 178
    assert bench.spmv_bytes(512 ** 3) == 45634027520


def test_vcycle_compulsory_bytes():
    # per row: 364 B zero-guess SymGS + 43 B residual/restriction + 4 B
    # prolongation + 667.5 B post-smoothing on every level above the coarsest,
    # 364 B on the coarsest
    assert 364 + 43 + 4 + 667.5 == 1078.5
    assert bench.vcycle_compulsory_bytes([1000]) == 364000.0
    assert bench.vcycle_compulsory_bytes([8, 1]) == 8 * 1078.5 + 364
    rows = [512 ** 3 // 8 ** k for k in range(4)]
    assert bench.vcycle_compulsory_bytes(rows) == 165205245952


def test_committed_traffic_matches_the_model():
    """The HBM bytes per SpMV launch read from the committed rocprofv3
    summaries lie just above the algorithmic bytes of that grid (x gathers
    mostly from cache), never below them."""
    for n in (320, 512):
        t = bench.profiled_traffic((n, n, n))
        assert t is not None, n
        ratio = t / bench.spmv_bytes(n ** 3)
        assert 1.0 <= ratio <= 1.15, (n, ratio)
    assert bench.profiled_traffic((48, 40, 32)) is None
