"""Host-only check of the batched launches' segment tables on the 64-problem
workload (nx, ny, nz from {16, ..., 96}, splitmix64 seed 7): every launch
the planner of rochpcg_amd/csrc/batch.hip builds for one CG iteration and its
preamble, with each per-level shape the tuner may pick.  The checker
(tests/cpp/batch_plan_test.cpp) runs the planner's block assignment, the
blocks' segment search and the XCD block order on the host and verifies that
every row of every segment is visited exactly once."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _colour_tables(oracle_mod, dims, nlevels=4):
    """Per level (sizes, offsets, m, mc) of a problem, as the device colouring
    leaves them: the parity classes (m/8 rows each) on even grids, the
    oracle's JPL colouring elsewhere."""
    out = []
    d = list(dims)
    for lv in range(nlevels):
        m = d[0] * d[1] * d[2]
        last = lv + 1 == nlevels
        mc = 0 if last else (d[0] // 2) * (d[1] // 2) * (d[2] // 2)
        if all(v % 2 == 0 for v in d):
            sizes = np.full(8, m // 8)
        else:
            O = oracle_mod.Problem(*d, nlevels=1, order=oracle_mod.ORDER_JPL)
            sizes = O.export(0)["sizes"]
            O.close()
        offsets = np.concatenate([[0], np.cumsum(sizes)])
        out.append((sizes, offsets, m, mc))
        d = [v // 2 for v in d]
    return out


def _launches(tables, block, rows):
    """(block, vector_kind, rows, [(off, size)]) of every launch of the
    preamble and one iteration, in the planner's order."""
    L = []
    m0 = [t[0][2] for t in tables]

    def vec():
        L.append((256, 1, 1, [(0, m) for m in m0]))

    def row(blk, r, segs):
        if segs:
            L.append((blk, 0, r, segs))
    vec()                                          # p = x
    row(block[0], 1, [(0, m) for m in m0])         # Ap = A p
    vec()                                          # r = b - Ap
    vec()                                          # ||r||
    nlev = len(tables[0])

    def mg(lv):
        C = max(len(t[lv][0]) for t in tables)
        last = lv + 1 == nlev

        def seg(t, c):
            sizes, offsets, _, _ = t[lv]
            return (int(offsets[c]), int(sizes[c])) if c < len(sizes) else None
        for c in range(C):
            row(block[lv], rows[lv], [s for s in (seg(t, c) for t in tables) if s])
        for c in range(C - 2, -1, -1):
            row(block[lv], rows[lv], [seg(t, c) for t in tables if c <= len(t[lv][0]) - 2])
        if last:
            return
        row(block[lv], 1, [(0, t[lv][3]) for t in tables])   # restriction
        mg(lv + 1)
        row(block[lv], 1, [(0, t[lv][3]) for t in tables])   # prolongation
        for c in range(C):
            row(block[lv], rows[lv], [s for s in (seg(t, c) for t in tables) if s])
        for c in range(C - 2, -1, -1):
            row(block[lv], rows[lv], [seg(t, c) for t in tables if c <= len(t[lv][0]) - 2])
    mg(0)
    vec()                                          # r'z
    vec()                                          # p update
    row(block[0], 1, [(0, m) for m in m0])         # Ap = A p (+ p'Ap)
    vec()                                          # r, x update and ||r||
    return L


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("bpt") / "batch_plan_test")
    r = subprocess.run(["/opt/rocm/bin/hipcc", "-std=c++17", "-O1", "--offload-arch=gfx950", "-x", "hip",
                        "-I" + os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "cpp", "batch_plan_test.cpp"),
                        "-o", exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_batch_plan_64_problems(oracle_mod, checker):
    from rochpcg_amd import vv
    dims = vv.batched_dims(64, seed=7)
    tables = [_colour_tables(oracle_mod, tuple(int(v) for v in d)) for d in dims]
    lines = []
    for block in (64, 128, 256):
        for rows in (1, 2):
            for blk, vk, r, segs in _launches(tables, [block] * 4, [rows] * 4):
                lines.append(f"{blk} {vk} {r} {len(segs)}")
                lines += [f"{o} {s}" for o, s in segs]
    # the shape the tuner settled on for this workload: 128 threads, two rows on the coarsest level
    for blk, vk, r, segs in _launches(tables, [128] * 4, [1, 1, 1, 2]):
        lines.append(f"{blk} {vk} {r} {len(segs)}")
        lines += [f"{o} {s}" for o, s in segs]
    out = subprocess.run([checker], input="\n".join(lines) + "\n", capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.startswith("ok ")
    n = int(out.stdout.split()[1])
    assert n > 7 * 100


def test_batch_plan_odd_segments(checker):
    """Segments of odd offset or size force one row per thread; single-row
    and single-element segments, and the reduction's block cap."""
    lines = ["256 0 2 3", "0 7", "7 2", "9 1",
             "64 0 2 2", "0 4096", "4096 4096",
             "256 1 1 2", "0 5000000", "0 1"]
    out = subprocess.run([checker], input="\n".join(lines) + "\n", capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.split()[:2] == ["ok", "3"]
