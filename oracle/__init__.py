"""TEST INFRASTRUCTURE ONLY -- ctypes wrapper of the CPU oracle.

The oracle (``hpcg_oracle.c``) restates the HPCG reference path of ROCm/rocHPCG
on the CPU; ``ref`` wraps ``_ref/libhpcg_ref.so``, the reference's own ``_ref``
sources compiled by ``make -C oracle ref``.  Only tests, ``__graft_entry__.smoke``
and the CPU-baseline leg of ``bench.py`` use this package, always as the
checker; the product (``rochpcg_amd``) never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
W = 27

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")

ORDER_NATURAL = 0
ORDER_JPL = 1
ORDER_CLOSED_FORM = 2


def build(ref: bool = True) -> None:
    """Compile liboracle.so (and oracle/_ref when the reference tree exists)."""
    targets = ["oracle"]
    if ref and os.path.isdir("/root/reference/src"):
        targets.append("ref")
    subprocess.run(["make", "-s", "-C", HERE] + targets, check=True)


def _load(name):
    path = os.path.join(HERE, name)
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is not built (run make -C oracle)")
    return C.CDLL(path)


_lib = None


def lib():
    global _lib
    if _lib is None:
        L = _load("liboracle.so")
        vp = C.c_void_p
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_get_threads.restype = C.c_int
        L.orc_create.restype = vp
        L.orc_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_destroy.argtypes = [vp]
        L.orc_nlevels.argtypes = [vp]
        L.orc_level_info.restype = C.c_int64
        L.orc_level_info.argtypes = [vp, C.c_int, _i32p]
        L.orc_export_level.argtypes = [vp, C.c_int] + [vp] * 9
        L.orc_generate.argtypes = [C.c_int, C.c_int, C.c_int, _i32p, _f64p, _u8p, _i32p, vp, vp]
        L.orc_f2c.argtypes = [C.c_int, C.c_int, C.c_int, _i32p]
        L.orc_row_hash.argtypes = [C.c_int, C.c_int, C.c_int, _i32p]
        L.orc_jpl.restype = C.c_int
        L.orc_jpl.argtypes = [C.c_int, _i32p, _i32p, C.c_uint, _i32p, _i32p]
        L.orc_closed_form_colors.argtypes = [C.c_int, C.c_int, C.c_int, _i32p]
        L.orc_perm_from_colors.argtypes = [C.c_int, _i32p, C.c_int, _i32p, _i32p, _i32p]
        L.orc_permute_matrix.argtypes = [C.c_int, _i32p, _f64p, _u8p, _i32p, _i32p, _f64p, _u8p, _i32p]
        L.orc_rhs_ones.argtypes = [vp, _f64p]
        L.orc_rhs_random.argtypes = [vp, C.c_uint64, _f64p, _f64p]
        L.orc_uniform.restype = C.c_double
        L.orc_uniform.argtypes = [C.c_uint64, C.c_int64]
        L.orc_permute_vector.argtypes = [vp, C.c_int, _f64p, _f64p]
        L.orc_unpermute_vector.argtypes = [vp, C.c_int, _f64p, _f64p]
        L.orc_spmv.argtypes = [vp, C.c_int, _f64p, _f64p]
        L.orc_symgs.argtypes = [vp, C.c_int, _f64p, _f64p]
        L.orc_mg.argtypes = [vp, C.c_int, _f64p, _f64p]
        L.orc_set_dot_mode.argtypes = [C.c_int]
        L.orc_dot.restype = C.c_double
        L.orc_dot.argtypes = [C.c_int, vp, vp]
        L.orc_waxpby.argtypes = [C.c_int, C.c_double, _f64p, C.c_double, _f64p, _f64p]
        L.orc_get_level_vectors.argtypes = [vp, C.c_int, vp, vp, vp]
        L.orc_cg.argtypes = [vp, _f64p, _f64p, C.c_int, C.c_double, C.c_int,
                             C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_double),
                             _f64p, _f64p]
        L.orc_copy_diagonal.argtypes = [vp, C.c_int, _f64p]
        L.orc_replace_diagonal.argtypes = [vp, C.c_int, _f64p]
        _lib = L
    return _lib


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


DOT_REFERENCE = 0      # the reference's OpenMP reduction
DOT_COMPENSATED = 1    # Neumaier-compensated sequential sum of the same products


def set_dot_mode(mode: int) -> None:
    """Summation of every dot product of the oracle (orc_dot, used by its CG)."""
    lib().orc_set_dot_mode(int(mode))


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Level:
    """Arrays of one level in the problem's row order (row-major [row, 27])."""

    def __init__(self, nx, ny, nz, m, mc, ncolors, nnz_total):
        self.nx, self.ny, self.nz, self.m, self.mc = nx, ny, nz, m, mc
        self.ncolors, self.nnz_total = ncolors, nnz_total


class Problem:
    """A 4-level (by default) HPCG hierarchy held by the CPU oracle."""

    def __init__(self, nx, ny=None, nz=None, nlevels=4, order=ORDER_JPL):
        ny = nx if ny is None else ny
        nz = nx if nz is None else nz
        self.h = lib().orc_create(nx, ny, nz, nlevels, order)
        if not self.h:
            raise ValueError(f"oracle cannot build {nx}x{ny}x{nz} with {nlevels} levels")
        self.nlevels = nlevels
        self.order = order
        self.levels = []
        for lv in range(nlevels):
            info = np.zeros(7, np.int32)
            nnz = lib().orc_level_info(self.h, lv, info)
            self.levels.append(Level(*[int(v) for v in info[:6]], int(nnz)))
        self.m = self.levels[0].m

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def export(self, level=0):
        L = self.levels[level]
        m = L.m
        out = dict(
            cols=np.empty((m, W), np.int32), vals=np.empty((m, W)), nnz=np.empty(m, np.uint8),
            diag=np.empty(m, np.int32), perm=np.empty(m, np.int32),
            f2c=np.empty(max(L.mc, 1), np.int32), sizes=np.empty(L.ncolors, np.int32),
            offsets=np.empty(L.ncolors + 1, np.int32), colors=np.empty(m, np.int32))
        lib().orc_export_level(self.h, level, *[_ptr(out[k]) for k in (
            "cols", "vals", "nnz", "diag", "perm", "f2c", "sizes", "offsets", "colors")])
        out["f2c"] = out["f2c"][:L.mc]
        return out

    # vectors -------------------------------------------------------------
    def rhs_ones(self):
        b = np.empty(self.m)
        lib().orc_rhs_ones(self.h, b)
        return b

    def rhs_random(self, seed):
        xs, b = np.empty(self.m), np.empty(self.m)
        lib().orc_rhs_random(self.h, seed, xs, b)
        return xs, b

    def permute(self, v_nat, level=0):
        out = np.empty(self.levels[level].m)
        lib().orc_permute_vector(self.h, level, np.ascontiguousarray(v_nat, np.float64), out)
        return out

    def unpermute(self, v, level=0):
        out = np.empty(self.levels[level].m)
        lib().orc_unpermute_vector(self.h, level, np.ascontiguousarray(v, np.float64), out)
        return out

    # kernels ---------------------------------------------------------------
    def spmv(self, x, level=0):
        y = np.empty(self.levels[level].m)
        lib().orc_spmv(self.h, level, np.ascontiguousarray(x, np.float64), y)
        return y

    def symgs(self, r, x, level=0):
        """One symmetric Gauss-Seidel step in place on x (returns x)."""
        lib().orc_symgs(self.h, level, np.ascontiguousarray(r, np.float64), x)
        return x

    def mg(self, r, level=0):
        x = np.empty(self.levels[level].m)
        lib().orc_mg(self.h, level, np.ascontiguousarray(r, np.float64), x)
        return x

    def level_vectors(self, level):
        L = self.levels[level]
        Axf, rc, xc = np.empty(L.m), np.empty(max(L.mc, 1)), np.empty(max(L.mc, 1))
        lib().orc_get_level_vectors(self.h, level, _ptr(Axf), _ptr(rc), _ptr(xc))
        return Axf, rc[:L.mc], xc[:L.mc]

    def cg(self, b, x0=None, max_iter=50, tol=0.0, precond=True):
        """CG_ref; returns dict(x, niters, normr, normr0, history, times)."""
        x = np.zeros(self.m) if x0 is None else np.array(x0, np.float64)
        hist = np.zeros(max_iter + 1)
        times = np.zeros(8)
        ni, nr, nr0 = C.c_int(0), C.c_double(0), C.c_double(0)
        lib().orc_cg(self.h, np.ascontiguousarray(b, np.float64), x, max_iter, tol,
                     1 if precond else 0, C.byref(ni), C.byref(nr), C.byref(nr0), hist, times)
        return dict(x=x, niters=ni.value, normr=nr.value, normr0=nr0.value,
                    history=hist[:ni.value + 1].copy(), times=times)

    def copy_diagonal(self, level=0):
        d = np.empty(self.levels[level].m)
        lib().orc_copy_diagonal(self.h, level, d)
        return d

    def replace_diagonal(self, d, level=0):
        lib().orc_replace_diagonal(self.h, level, np.ascontiguousarray(d, np.float64))


def dot(x, y):
    x = np.ascontiguousarray(x, np.float64)
    y = x if y is x else np.ascontiguousarray(y, np.float64)
    return lib().orc_dot(len(x), x.ctypes.data_as(C.c_void_p), y.ctypes.data_as(C.c_void_p))


def waxpby(alpha, x, beta, y):
    w = np.empty(len(x))
    lib().orc_waxpby(len(x), alpha, np.ascontiguousarray(x, np.float64), beta,
                     np.ascontiguousarray(y, np.float64), w)
    return w


def generate(nx, ny, nz):
    m = nx * ny * nz
    cols, vals = np.empty((m, W), np.int32), np.empty((m, W))
    nnz, diag = np.empty(m, np.uint8), np.empty(m, np.int32)
    b, xe = np.empty(m), np.empty(m)
    lib().orc_generate(nx, ny, nz, cols, vals, nnz, diag, _ptr(b), _ptr(xe))
    return dict(cols=cols, vals=vals, nnz=nnz, diag=diag, b=b, xexact=xe)


def f2c(nxf, nyf, nzf):
    out = np.empty((nxf // 2) * (nyf // 2) * (nzf // 2), np.int32)
    lib().orc_f2c(nxf, nyf, nzf, out)
    return out


def row_hash(nx, ny, nz):
    out = np.empty(nx * ny * nz, np.int32)
    lib().orc_row_hash(nx, ny, nz, out)
    return out


def jpl(cols, hash_, seed=0x586744):
    m = len(hash_)
    colors = np.empty(m, np.int32)
    rs = np.zeros(128, np.int32)
    n = lib().orc_jpl(m, np.ascontiguousarray(cols, np.int32), np.ascontiguousarray(hash_, np.int32),
                      seed, colors, rs)
    return colors, n, rs


def closed_form_colors(nx, ny, nz):
    out = np.empty(nx * ny * nz, np.int32)
    lib().orc_closed_form_colors(nx, ny, nz, out)
    return out


def perm_from_colors(colors, ncolors):
    m = len(colors)
    perm = np.empty(m, np.int32)
    sizes, offsets = np.empty(ncolors, np.int32), np.empty(ncolors + 1, np.int32)
    rc = lib().orc_perm_from_colors(m, np.ascontiguousarray(colors, np.int32), ncolors, perm,
                                    sizes, offsets)
    if rc:
        raise ValueError("colour out of range")
    return perm, sizes, offsets


def permute_matrix(cols, vals, nnz, perm):
    m = len(nnz)
    pc, pv = np.empty((m, W), np.int32), np.empty((m, W))
    pn, pd = np.empty(m, np.uint8), np.empty(m, np.int32)
    rc = lib().orc_permute_matrix(m, cols, vals, nnz, np.ascontiguousarray(perm, np.int32),
                                  pc, pv, pn, pd)
    if rc:
        raise ValueError("permuted row lost its diagonal")
    return pc, pv, pn, pd


def uniform(seed, i):
    return lib().orc_uniform(seed, i)


def hpcg_flops_per_iter(dims, levels=4):
    """HPCG flop count of one steady-state CG iteration (ReportResults.cpp:73-89)."""
    nx, ny, nz = dims
    nrow = nx * ny * nz
    fl = 6.0 * nrow + 6.0 * nrow  # 3 ddot + 3 waxpby, 2 flops each
    dd = list(dims)
    nnz = [(3 * dd[0] - 2) * (3 * dd[1] - 2) * (3 * dd[2] - 2)]
    for _ in range(levels - 1):
        dd = [d // 2 for d in dd]
        nnz.append((3 * dd[0] - 2) * (3 * dd[1] - 2) * (3 * dd[2] - 2))
    fl += 2.0 * nnz[0]  # SpMV
    for lv in range(levels - 1):
        fl += 4.0 * nnz[lv] + 2.0 * nnz[lv] + 4.0 * nnz[lv]
    fl += 4.0 * nnz[levels - 1]
    return fl


# ---------------------------------------------------------------------------
# numpy restatements for full-size checks (no matrix is stored)

def stencil_apply(x, nx, ny, nz, chunk=32):
    """y = A x for the 27-point operator of GenerateProblem_ref.cpp:131-182 in
    natural order (diagonal 26, in-range neighbours -1): y = 27 x - (3x3x3
    box sum of x, zero outside the grid).  Sums in another order than
    ComputeSPMV_ref, so it agrees with it to rounding, not bit for bit.
    Works through z-slabs of `chunk` planes to bound host memory."""
    X = np.ascontiguousarray(x, np.float64).reshape(nz, ny, nx)
    y = np.empty_like(X)
    for z0 in range(0, nz, chunk):
        z1 = min(nz, z0 + chunk)
        lo, hi = max(0, z0 - 1), min(nz, z1 + 1)
        # P[k] holds plane z0 - 1 + k (zero outside the grid)
        P = np.zeros((z1 - z0 + 2, ny + 2, nx + 2))
        P[lo - z0 + 1:hi - z0 + 1, 1:-1, 1:-1] = X[lo:hi]
        S = P[:, :, :-2] + P[:, :, 1:-1] + P[:, :, 2:]
        S = S[:, :-2] + S[:, 1:-1] + S[:, 2:]
        S = S[:-2] + S[1:-1] + S[2:]
        y[z0:z1] = 27.0 * X[z0:z1] - S
    return y.ravel()


def uniform_vector(seed, m):
    """u_i = orc_uniform(seed, i) for i < m (splitmix64, vectorised)."""
    i = np.arange(1, m + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + i * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def split_ell_rows(nx, ny, nz, perm, iperm, rows):
    """Expected multicolour ELL entries of multicolour rows `rows` in the
    split-padded layout: the row's entries (perm of its natural neighbours,
    ascending, as Permute.cpp:95-149 sorts them) with the lower entries and the
    diagonal from slot 0 and the upper entries ending at slot 26, -1 / 0.0
    between.  Returns cols[len(rows), 27] (int32) and vals (float64)."""
    rows = np.asarray(rows, np.int64)
    nat = iperm[rows].astype(np.int64)
    nxy = nx * ny
    iz, rem = nat // nxy, nat % nxy
    iy, ix = rem // nx, rem % nx
    nb = np.full((len(rows), 27), np.iinfo(np.int64).max, np.int64)
    k = 0
    for sz in (-1, 0, 1):
        for sy in (-1, 0, 1):
            for sx in (-1, 0, 1):
                z, y, xx = iz + sz, iy + sy, ix + sx
                ok = (z >= 0) & (z < nz) & (y >= 0) & (y < ny) & (xx >= 0) & (xx < nx)
                j = np.where(ok, z * nxy + y * nx + xx, 0)
                nb[:, k] = np.where(ok, perm[j].astype(np.int64), np.iinfo(np.int64).max)
                k += 1
    nb.sort(axis=1)
    cnt = (nb != np.iinfo(np.int64).max).sum(axis=1)
    d = (nb < rows[:, None]).sum(axis=1)            # lower entries
    cols = np.full((len(rows), 27), -1, np.int64)
    vals = np.zeros((len(rows), 27))
    slot = np.arange(27)[None, :]
    lower_or_diag = slot <= d[:, None]
    cols[lower_or_diag] = nb[lower_or_diag]
    nup = cnt - d - 1
    us = 27 - nup
    upper = slot >= us[:, None]
    # upper entry q of a row (q = slot - us) is its sorted entry d + 1 + q
    src = np.clip(d[:, None] + 1 + (slot - us[:, None]), 0, 26)
    cols[upper] = np.take_along_axis(nb, src, 1)[upper]
    vals[cols >= 0] = -1.0
    vals[slot == d[:, None]] = 26.0
    return cols.astype(np.int32), vals


# ---------------------------------------------------------------------------
_ref = None


def ref_available():
    return os.path.exists(os.path.join(HERE, "_ref", "libhpcg_ref.so"))


def reflib():
    """The reference's own _ref sources built into oracle/_ref (checker only)."""
    global _ref
    if _ref is None:
        L = _load(os.path.join("_ref", "libhpcg_ref.so"))
        vp = C.c_void_p
        L.ref_generate.argtypes = [C.c_int, C.c_int, C.c_int, _i32p, _f64p, _u8p, _i32p,
                                   _f64p, _f64p, _f64p]
        L.ref_wrap.restype = vp
        L.ref_wrap.argtypes = [C.c_int, _i32p, vp, vp, vp, vp, vp]
        L.ref_destroy.argtypes = [vp]
        L.ref_spmv.argtypes = [vp, C.c_int, _f64p, _f64p]
        L.ref_symgs.argtypes = [vp, C.c_int, _f64p, _f64p]
        L.ref_mg.argtypes = [vp, C.c_int, _f64p, _f64p]
        L.ref_dot.restype = C.c_double
        L.ref_dot.argtypes = [C.c_int, vp, vp]
        L.ref_waxpby.argtypes = [C.c_int, C.c_double, _f64p, C.c_double, _f64p, _f64p]
        L.ref_cg.argtypes = [vp, _f64p, _f64p, C.c_int, C.c_double, C.c_int,
                             C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_double),
                             _f64p]
        _ref = L
    return _ref


def ref_generate(nx, ny, nz):
    m = nx * ny * nz
    out = dict(cols=np.empty((m, W), np.int32), vals=np.empty((m, W)), nnz=np.empty(m, np.uint8),
               diag=np.empty(m, np.int32), b=np.empty(m), x=np.empty(m), xexact=np.empty(m))
    reflib().ref_generate(nx, ny, nz, out["cols"], out["vals"], out["nnz"], out["diag"],
                          out["b"], out["x"], out["xexact"])
    return out


class RefHierarchy:
    """The reference's _ref kernels over the same arrays as an oracle Problem."""

    def __init__(self, prob: Problem):
        n = prob.nlevels
        self.exports = [prob.export(lv) for lv in range(n)]
        dims = np.array([[L.nx, L.ny, L.nz] for L in prob.levels], np.int32).ravel()

        def arr(key):
            a = (C.c_void_p * n)()
            for i, e in enumerate(self.exports):
                a[i] = e[key].ctypes.data
            return a
        self._keep = [arr(k) for k in ("cols", "vals", "nnz", "diag", "f2c")]
        self.h = reflib().ref_wrap(n, dims, *self._keep)
        self.m = prob.m
        self.ms = [L.m for L in prob.levels]

    def close(self):
        if self.h:
            reflib().ref_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def spmv(self, x, level=0):
        y = np.empty(self.ms[level])
        reflib().ref_spmv(self.h, level, np.array(x, np.float64), y)
        return y

    def symgs(self, r, x, level=0):
        reflib().ref_symgs(self.h, level, np.array(r, np.float64), x)
        return x

    def mg(self, r, level=0):
        x = np.empty(self.ms[level])
        reflib().ref_mg(self.h, level, np.array(r, np.float64), x)
        return x

    def cg(self, b, max_iter=50, tol=0.0, precond=True):
        x = np.zeros(self.m)
        times = np.zeros(8)
        ni, nr, nr0 = C.c_int(0), C.c_double(0), C.c_double(0)
        reflib().ref_cg(self.h, np.array(b, np.float64), x, max_iter, tol, 1 if precond else 0,
                        C.byref(ni), C.byref(nr), C.byref(nr0), times)
        return dict(x=x, niters=ni.value, normr=nr.value, normr0=nr0.value)

    def history(self, b, max_iter=50, precond=True):
        """normr after k iterations for k = 0..max_iter (CG_ref restarted per k)."""
        out = []
        for k in range(max_iter + 1):
            r = self.cg(b, max_iter=k, tol=0.0, precond=precond)
            out.append(r["normr"] if k else r["normr0"])
        return np.array(out)
