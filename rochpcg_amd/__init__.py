"""rochpcg_amd -- MI355X-native (gfx950) HPCG solver hot path.

Python host binding of ``librochpcg_amd.so`` (the C ABI declared in
``include/rochpcg_amd.h``).  It mirrors the reference's optimized-path
interface of ROCm/rocHPCG (``ComputeSPMV``, ``ComputeSYMGS``, ``ComputeMG``,
``ComputeDotProduct``, ``ComputeWAXPBY``, ``CG``; src/ComputeSPMV.hpp:20 and
siblings) on device buffers in multicolour row order.

There is no CPU fallback: importing fails loudly when the HIP library is not
built, and every call raises ``HPCGError`` when the library reports an error.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librochpcg_amd.so")

STORE_F64 = 1
STORE_MIXED = 2
ELL_SLICE = 128   # rows per slice of the sliced ELL (HPCGX_ELL_SLICE)
MAX_COLORS = 128  # colours a level may hold (HPCGX_MAX_COLORS)
CG_F64 = 0
CG_MIXED = 1

# hpcgx_time_operator op codes
(OP_SPMV, OP_SYMGS, OP_SYMGS_ZERO, OP_MG, OP_FWD_ZERO, OP_BWD_ZERO, OP_FWD_FULL, OP_BWD_FULL,
 OP_RESTRICT, OP_READ_STREAM) = range(10)
PHASES = ("spmv", "mg", "ddot", "waxpby", "true_residual")
# launch-shape operators (hpcgx_set_shape)
(SH_SPMV, SH_FWD_ZERO, SH_BWD_ZERO, SH_FWD_FULL, SH_BWD_FULL, SH_RESTRICT,
 SH_FWD_ZERO_N, SH_BWD_ZERO_N, SH_FWD_FULL_N, SH_BWD_FULL_N) = range(10)


class HPCGError(RuntimeError):
    pass


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built; run `python -c 'import __graft_entry__ as g; g.build()'` "
            "(rochpcg_amd has no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    vp, i64, i32, f64 = C.c_void_p, C.c_int64, C.c_int, C.c_double
    P = C.POINTER
    sig = {
        "hpcgx_problem_create": ([i32, i32, i32, i32, i32, vp, P(vp)], i32),
        "hpcgx_problem_destroy": ([vp], i32),
        "hpcgx_level_info": ([vp, i32, P(i64)], i32),
        "hpcgx_color_table": ([vp, i32, vp, vp, vp], i32),
        "hpcgx_level_arrays": ([vp, i32, P(vp)], i32),
        "hpcgx_vectors": ([vp, P(vp)], i32),
        "hpcgx_make_rhs": ([vp, i32, C.c_uint64], i32),
        "hpcgx_check_problem": ([vp, i32, i32, P(i64)], i32),
        "hpcgx_spmv": ([vp, i32, vp, vp], i32),
        "hpcgx_symgs": ([vp, i32, vp, vp], i32),
        "hpcgx_symgs_zero": ([vp, i32, vp, vp], i32),
        "hpcgx_mg": ([vp, i32, vp, vp], i32),
        "hpcgx_mg_mixed": ([vp, vp, vp], i32),
        "hpcgx_restriction": ([vp, i32, vp, vp, P(vp)], i32),
        "hpcgx_prolongation": ([vp, i32, vp, P(vp)], i32),
        "hpcgx_spmv_injected": ([vp, i32, vp, vp], i32),
        "hpcgx_restriction_axf": ([vp, i32, vp, vp, P(vp)], i32),
        "hpcgx_residual": ([vp, i64, vp, vp, P(f64)], i32),
        "hpcgx_dot": ([vp, i64, vp, vp, P(f64)], i32),
        "hpcgx_dot_noalloc": ([i64, vp, vp, P(f64), vp], i32),
        "hpcgx_waxpby": ([vp, i64, f64, vp, f64, vp, vp], i32),
        "hpcgx_fused_waxpby_dot": ([vp, i64, f64, vp, vp, P(f64)], i32),
        "hpcgx_copy_diagonal": ([vp, i32, vp], i32),
        "hpcgx_replace_diagonal": ([vp, i32, vp], i32),
        "hpcgx_testcg_scale": ([vp, vp, vp], i32),
        "hpcgx_cg": ([vp, i32, vp, vp, i32, f64, i32, i32, P(i32), P(f64), P(f64), vp, i32, vp], i32),
        "hpcgx_cg_begin": ([vp, i32, vp, vp, i32, f64, i32], i32),
        "hpcgx_cg_iterate": ([vp, i32, i32], i32),
        "hpcgx_cg_iterate_profiled": ([vp, i32, vp], i32),
        "hpcgx_cg_end": ([vp, P(i32), P(f64), P(f64), vp, i32], i32),
        "hpcgx_set_residual_replacement": ([vp, i32], i32),
        "hpcgx_spmv_ref": ([vp, i32, vp, vp], i32),
        "hpcgx_symgs_ref": ([vp, i32, vp, vp], i32),
        "hpcgx_mg_ref": ([vp, i32, vp, vp], i32),
        "hpcgx_cg_ref": ([vp, vp, vp, i32, f64, i32, P(i32), P(f64), P(f64), vp, i32, vp], i32),
        "hpcgx_flops_per_iteration": ([vp], f64),
        "hpcgx_iteration_bytes": ([vp, i32, P(f64), P(f64)], i32),
        "hpcgx_time_operator": ([vp, i32, i32, i32, i32, P(f64)], i32),
        "hpcgx_set_shape": ([vp, i32, i32, i32, i32, i32], i32),
        "hpcgx_set_sweep_columns": ([vp, i32, i32], i32),
        "hpcgx_get_shape": ([vp, i32, i32, i32, P(i32), P(i32)], i32),
        "hpcgx_synchronize": ([vp], i32),
        "hpcgx_batch_create": ([i32, vp, i32, i32, vp, P(vp)], i32),
        "hpcgx_batch_destroy": ([vp], i32),
        "hpcgx_batch_problem": ([vp, i32, P(vp)], i32),
        "hpcgx_batch_set_block": ([vp, i32, i32], i32),
        "hpcgx_batch_set_rows": ([vp, i32, i32], i32),
        "hpcgx_batch_cg": ([vp, i32, i32, f64, i32, vp, vp, vp], i32),
        "hpcgx_batch_history": ([vp, i32, vp, i32, P(i32)], i32),
        "hpcgx_batch_profile": ([vp, i32, i32, vp, vp, P(f64)], i32),
        "hpcgx_malloc": ([P(vp), C.c_size_t], i32),
        "hpcgx_free": ([vp], i32),
        "hpcgx_memcpy_h2d": ([vp, vp, C.c_size_t], i32),
        "hpcgx_memcpy_d2h": ([vp, vp, C.c_size_t], i32),
        "hpcgx_memcpy_d2d": ([vp, vp, C.c_size_t], i32),
        "hpcgx_device_count": ([P(i32)], i32),
        "hpcgx_last_error": ([], C.c_char_p),
        "hpcgx_version": ([], C.c_char_p),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    return L


lib = _load()
EXPORTED = tuple(n for n in dir(lib) if n.startswith("hpcgx_"))


def _check(rc):
    if rc != 0:
        raise HPCGError(lib.hpcgx_last_error().decode() or f"status {rc}")


def dot_noalloc(x, y, n, stream=None) -> float:
    """x'y over n doubles without a problem handle or any device allocation
    (hpcgx_dot_noalloc: the reduction scratch is a fixed device array of the
    library, shared by every call on the GPU)."""
    res = C.c_double()
    _check(lib.hpcgx_dot_noalloc(int(n), _p(x), _p(y), C.byref(res), stream))
    return res.value


def device_count() -> int:
    n = C.c_int(0)
    rc = lib.hpcgx_device_count(C.byref(n))
    return n.value if rc == 0 else 0


class DeviceArray:
    """A raw device allocation with numpy-typed host transfers."""

    def __init__(self, n, dtype=np.float64, ptr=None, owner=True):
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        if ptr is None:
            p = C.c_void_p()
            _check(lib.hpcgx_malloc(C.byref(p), max(self.nbytes, 8)))
            self.ptr = p.value
            self.owner = True
        else:
            self.ptr = int(ptr)
            self.owner = owner

    @classmethod
    def from_host(cls, a, dtype=None):
        a = np.ascontiguousarray(a, dtype=dtype or np.asarray(a).dtype)
        d = cls(a.size, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if a.size != self.n:
            raise ValueError(f"size mismatch: {a.size} != {self.n}")
        _check(lib.hpcgx_memcpy_h2d(self.ptr, a.ctypes.data, self.nbytes))

    def host(self, start=0, count=None):
        """Elements [start, start + count) copied to a host array."""
        count = self.n - start if count is None else count
        if start < 0 or count < 0 or start + count > self.n:
            raise ValueError(f"range [{start}, {start + count}) outside an array of {self.n}")
        out = np.empty(count, self.dtype)
        if count:
            _check(lib.hpcgx_memcpy_d2h(out.ctypes.data, self.ptr + start * self.dtype.itemsize,
                                        count * self.dtype.itemsize))
        return out

    def free(self):
        if self.owner and self.ptr:
            lib.hpcgx_free(self.ptr)
        self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def ell_padded_rows(m):
    """Rows the sliced ELL arrays of an m-row level hold (m rounded up to a
    whole slice)."""
    return -(-int(m) // ELL_SLICE) * ELL_SLICE


def ell_unslice(flat, m):
    """A sliced ELL array (slice s holds slot p of its rows at
    [(s*27 + p)*ELL_SLICE, +ELL_SLICE)) as an (m, 27) array, row by row."""
    ns = ell_padded_rows(m) // ELL_SLICE
    return flat.reshape(ns, 27, ELL_SLICE).transpose(0, 2, 1).reshape(ns * ELL_SLICE, 27)[:m]


def ell_rows(view, r0, count):
    """Rows [r0, r0 + count) of a device ELL array view as a (count, 27)
    host array (reads only the slices that hold them)."""
    s0, s1 = r0 // ELL_SLICE, -(-(r0 + count) // ELL_SLICE)
    blk = view.host(s0 * 27 * ELL_SLICE, (s1 - s0) * 27 * ELL_SLICE)
    rows = blk.reshape(s1 - s0, 27, ELL_SLICE).transpose(0, 2, 1).reshape(-1, 27)
    a = r0 - s0 * ELL_SLICE
    return rows[a:a + count]


def _p(a):
    if a is None:
        return None
    return a.ptr if isinstance(a, DeviceArray) else int(a)


class Level:
    def __init__(self, info):
        (self.nx, self.ny, self.nz, self.m, self.nnz, self.ncolors, self.mc,
         self.stores) = [int(v) for v in info]


class Problem:
    """A multigrid HPCG hierarchy resident on the GPU (multicolour order).

    Setup = GenerateProblem + GenerateCoarseProblem + OptimizeProblem of the
    reference, done on the device.  b = A*ones, x = 0, xexact = 1 initially.
    """

    def __init__(self, nx, ny=None, nz=None, nlevels=4, stores=STORE_F64, handle=None):
        if handle is None:
            ny = nx if ny is None else ny
            nz = nx if nz is None else nz
            h = C.c_void_p()
            _check(lib.hpcgx_problem_create(nx, ny, nz, nlevels, stores, None, C.byref(h)))
            handle = h.value
        self.h = handle
        self.nlevels = nlevels
        self.levels = []
        for lv in range(nlevels):
            info = (C.c_int64 * 8)()
            _check(lib.hpcgx_level_info(self.h, lv, info))
            self.levels.append(Level(info))
        self.m = self.levels[0].m
        ptrs = (C.c_void_p * 8)()
        _check(lib.hpcgx_vectors(self.h, ptrs))
        names = ("b", "x", "xexact", "r", "z", "p", "Ap", "z32")
        self.vec = {}
        for name, ptr in zip(names, ptrs):
            if ptr:
                dt = np.float32 if name == "z32" else np.float64
                self.vec[name] = DeviceArray(self.m, dt, ptr=ptr, owner=False)

    def close(self):
        if getattr(self, "h", None):
            lib.hpcgx_problem_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- introspection -------------------------------------------------------
    def color_table(self, level=0):
        L = self.levels[level]
        sizes = np.zeros(L.ncolors, np.int32)
        offsets = np.zeros(L.ncolors + 1, np.int32)
        rounds = np.zeros(MAX_COLORS, np.int32)
        _check(lib.hpcgx_color_table(self.h, level, sizes.ctypes.data, offsets.ctypes.data,
                                     rounds.ctypes.data))
        return sizes, offsets, rounds

    def level_arrays(self, level=0):
        """Device arrays of a level copied to host numpy arrays; the ELL
        arrays as (rows, 27), row by row."""
        L = self.levels[level]
        ptrs = (C.c_void_p * 13)()
        _check(lib.hpcgx_level_arrays(self.h, level, ptrs))
        m, mc = L.m, L.mc
        me = ell_padded_rows(m)

        def get(ptr, n, dt):
            if not ptr or n == 0:
                return None
            return DeviceArray(n, dt, ptr=ptr, owner=False).host()

        def ell(ptr, dt):
            a = get(ptr, 27 * me, dt)
            return None if a is None else ell_unslice(a, m)
        out = dict(
            ell_col=ell(ptrs[0], np.int32), ell_val=ell(ptrs[1], np.float64),
            ell_f16=ell(ptrs[2], np.float16), diag_idx=get(ptrs[3], m, np.int32),
            inv_diag=get(ptrs[4], m, np.float64), perm=get(ptrs[5], m, np.int32),
            colors=get(ptrs[6], m, np.int32), f2c=get(ptrs[7], mc, np.int32),
            f2c_fine=get(ptrs[8], mc, np.int32), f2c_coarse=get(ptrs[9], mc, np.int32),
            desc=get(ptrs[10], m, np.int32), diag=get(ptrs[11], m, np.float64),
            diag_f16=get(ptrs[12], m, np.float16))
        return out

    def level_views(self, level=0):
        """Device arrays of a level as non-owning DeviceArray views (no copy);
        the ELL arrays are sliced (read rows with ell_rows)."""
        L = self.levels[level]
        ptrs = (C.c_void_p * 13)()
        _check(lib.hpcgx_level_arrays(self.h, level, ptrs))
        m, mc = L.m, L.mc
        me = ell_padded_rows(m)
        spec = (("ell_col", 27 * me, np.int32), ("ell_val", 27 * me, np.float64),
                ("ell_f16", 27 * me, np.float16), ("diag_idx", m, np.int32), ("inv_diag", m, np.float64),
                ("perm", m, np.int32), ("colors", m, np.int32), ("f2c", mc, np.int32),
                ("f2c_fine", mc, np.int32), ("f2c_coarse", mc, np.int32), ("desc", m, np.int32),
                ("diag", m, np.float64), ("diag_f16", m, np.float16))
        return {name: DeviceArray(n, dt, ptr=ptr, owner=False)
                for (name, n, dt), ptr in zip(spec, ptrs) if ptr and n}

    def make_rhs(self, kind=0, seed=0):
        _check(lib.hpcgx_make_rhs(self.h, kind, seed))

    CHECK_FIELDS = ("perm", "color", "ell", "value", "desc", "vectors", "nnz", "rows")

    def check_problem(self, level=0, vectors=False):
        """CheckProblem on the device (hpcgx_check_problem): error counts per
        kind, the stored nonzeros and the rows checked."""
        counts = (C.c_int64 * 8)()
        _check(lib.hpcgx_check_problem(self.h, level, int(vectors), counts))
        return dict(zip(self.CHECK_FIELDS, (int(v) for v in counts)))

    # -- operators (reference names in the C ABI docs) -------------------------
    def spmv(self, x, y, level=0):
        _check(lib.hpcgx_spmv(self.h, level, _p(x), _p(y)))

    def symgs(self, r, x, level=0):
        _check(lib.hpcgx_symgs(self.h, level, _p(r), _p(x)))

    def symgs_zero(self, r, x, level=0):
        _check(lib.hpcgx_symgs_zero(self.h, level, _p(r), _p(x)))

    def mg(self, r, x, level=0):
        _check(lib.hpcgx_mg(self.h, level, _p(r), _p(x)))

    def mg_mixed(self, r, x32):
        _check(lib.hpcgx_mg_mixed(self.h, _p(r), _p(x32)))

    def restriction(self, rf, xf, level=0):
        out = C.c_void_p()
        _check(lib.hpcgx_restriction(self.h, level, _p(rf), _p(xf), C.byref(out)))
        return DeviceArray(self.levels[level + 1].m, np.float64, ptr=out.value, owner=False)

    def spmv_injected(self, xf, Axf, level=0):
        """ComputeSPMV into MGData::Axf: A xf at the injected rows only."""
        _check(lib.hpcgx_spmv_injected(self.h, level, _p(xf), _p(Axf)))

    def restriction_axf(self, rf, Axf, level=0):
        """ComputeRestriction (reference mode): rc = rf - Axf at the injected rows."""
        out = C.c_void_p()
        _check(lib.hpcgx_restriction_axf(self.h, level, _p(rf), _p(Axf), C.byref(out)))
        return DeviceArray(self.levels[level + 1].m, np.float64, ptr=out.value, owner=False)

    def residual(self, v1, v2, n=None):
        """ComputeResidual: max_i |v1_i - v2_i| over the first n entries."""
        res = C.c_double()
        n = self.m if n is None else n
        _check(lib.hpcgx_residual(self.h, n, _p(v1), _p(v2), C.byref(res)))
        return res.value

    def prolongation(self, xf, level=0):
        out = C.c_void_p()
        _check(lib.hpcgx_prolongation(self.h, level, _p(xf), C.byref(out)))
        return DeviceArray(self.levels[level + 1].m, np.float64, ptr=out.value, owner=False)

    def dot(self, x, y, n=None):
        res = C.c_double()
        n = self.m if n is None else n
        _check(lib.hpcgx_dot(self.h, n, _p(x), _p(y), C.byref(res)))
        return res.value

    def waxpby(self, alpha, x, beta, y, w, n=None):
        n = self.m if n is None else n
        _check(lib.hpcgx_waxpby(self.h, n, alpha, _p(x), beta, _p(y), _p(w)))

    def fused_waxpby_dot(self, alpha, x, y, n=None):
        res = C.c_double()
        n = self.m if n is None else n
        _check(lib.hpcgx_fused_waxpby_dot(self.h, n, alpha, _p(x), _p(y), C.byref(res)))
        return res.value

    def copy_diagonal(self, d, level=0):
        _check(lib.hpcgx_copy_diagonal(self.h, level, _p(d)))

    def replace_diagonal(self, d, level=0):
        _check(lib.hpcgx_replace_diagonal(self.h, level, _p(d)))

    def testcg_scale(self, d, b):
        _check(lib.hpcgx_testcg_scale(self.h, _p(d), _p(b)))

    def set_shape(self, block, rows, level=-1, op=-1, mixed=False):
        _check(lib.hpcgx_set_shape(self.h, int(mixed), level, op, block, rows))

    def set_sweep_columns(self, columns, level=-1):
        """Block order of the SymGS colour launches: 0 chooses, 1 keeps row
        order, k > 1 sweeps k columns per plane (results do not change)."""
        _check(lib.hpcgx_set_sweep_columns(self.h, level, columns))

    def get_shape(self, level, op, mixed=False):
        b, r = C.c_int(), C.c_int()
        _check(lib.hpcgx_get_shape(self.h, int(mixed), level, op, C.byref(b), C.byref(r)))
        return b.value, r.value

    def synchronize(self):
        _check(lib.hpcgx_synchronize(self.h))

    # -- CG ------------------------------------------------------------------
    def cg(self, b=None, x=None, max_iter=50, tol=0.0, mode=CG_F64, precond=True, use_graph=True,
           history=True):
        """CG (CG.cpp:104-200); returns dict(niters, normr, normr0, history, seconds)."""
        ni, nr, nr0 = C.c_int(), C.c_double(), C.c_double()
        hist = np.zeros(max_iter + 1)
        times = np.zeros(8)
        _check(lib.hpcgx_cg(self.h, mode, _p(b), _p(x), max_iter, tol, int(precond), int(use_graph),
                            C.byref(ni), C.byref(nr), C.byref(nr0),
                            hist.ctypes.data if history else None, max_iter + 1, times.ctypes.data))
        return dict(niters=ni.value, normr=nr.value, normr0=nr0.value,
                    history=hist[:ni.value + 1].copy(), seconds=times[0])

    def cg_begin(self, mode=CG_F64, b=None, x=None, max_iter=50, tol=0.0, precond=True):
        _check(lib.hpcgx_cg_begin(self.h, mode, _p(b), _p(x), max_iter, tol, int(precond)))

    def cg_iterate(self, count, use_graph=True):
        _check(lib.hpcgx_cg_iterate(self.h, count, int(use_graph)))

    def cg_iterate_profiled(self, count):
        ms = np.zeros(len(PHASES))
        _check(lib.hpcgx_cg_iterate_profiled(self.h, count, ms.ctypes.data))
        return dict(zip(PHASES, ms))

    def cg_end(self, hist_len=0):
        ni, nr, nr0 = C.c_int(), C.c_double(), C.c_double()
        hist = np.zeros(max(hist_len, 1))
        _check(lib.hpcgx_cg_end(self.h, C.byref(ni), C.byref(nr), C.byref(nr0),
                                hist.ctypes.data if hist_len else None, hist_len))
        return dict(niters=ni.value, normr=nr.value, normr0=nr0.value,
                    history=hist[:min(hist_len, ni.value + 1)])

    # -- the reference's natural-order path on the device ----------------------
    def spmv_ref(self, x, y, level=0):
        """ComputeSPMV_ref in natural row order (device vectors)."""
        _check(lib.hpcgx_spmv_ref(self.h, level, _p(x), _p(y)))

    def symgs_ref(self, r, x, level=0):
        """ComputeSYMGS_ref: sequential forward + backward Gauss-Seidel in natural order."""
        _check(lib.hpcgx_symgs_ref(self.h, level, _p(r), _p(x)))

    def mg_ref(self, r, x, level=0):
        """ComputeMG_ref in natural order."""
        _check(lib.hpcgx_mg_ref(self.h, level, _p(r), _p(x)))

    def cg_ref(self, b=None, x=None, max_iter=50, tol=0.0, precond=True):
        """CG_ref in natural order on the device; b None = the problem's b in
        natural order, x None = a zeroed work vector.  Returns dict(niters,
        normr, normr0, history, seconds)."""
        ni, nr, nr0 = C.c_int(), C.c_double(), C.c_double()
        hist = np.zeros(max_iter + 1)
        times = np.zeros(8)
        _check(lib.hpcgx_cg_ref(self.h, _p(b), _p(x), max_iter, tol, int(precond), C.byref(ni),
                                C.byref(nr), C.byref(nr0), hist.ctypes.data, max_iter + 1,
                                times.ctypes.data))
        return dict(niters=ni.value, normr=nr.value, normr0=nr0.value,
                    history=hist[:ni.value + 1].copy(), seconds=times[0])

    def set_residual_replacement(self, every):
        _check(lib.hpcgx_set_residual_replacement(self.h, every))

    def flops_per_iteration(self):
        return lib.hpcgx_flops_per_iteration(self.h)

    def iteration_bytes(self, mode=CG_F64):
        """Algorithmic HBM bytes (V-cycle, whole CG iteration) of this library's kernels."""
        mg, it = C.c_double(), C.c_double()
        _check(lib.hpcgx_iteration_bytes(self.h, mode, C.byref(mg), C.byref(it)))
        return mg.value, it.value

    def time_operator(self, op, level=0, mixed=False, reps=10):
        ms = C.c_double()
        _check(lib.hpcgx_time_operator(self.h, op, level, int(mixed), reps, C.byref(ms)))
        return ms.value


def version() -> str:
    return lib.hpcgx_version().decode()


class Batch:
    """Independent problems of uneven sizes solved together (one launch per
    operator, level and colour across all problems)."""

    def __init__(self, dims, nlevels=4, stores=STORE_F64):
        dims = np.ascontiguousarray(np.asarray(dims, dtype=np.int32).reshape(-1, 3))
        h = C.c_void_p()
        _check(lib.hpcgx_batch_create(len(dims), dims.ctypes.data, nlevels, stores, None, C.byref(h)))
        self.h = h.value
        self.dims = dims
        self.problems = []
        for k in range(len(dims)):
            ph = C.c_void_p()
            _check(lib.hpcgx_batch_problem(self.h, k, C.byref(ph)))
            self.problems.append(Problem(0, nlevels=nlevels, handle=ph.value))

    def close(self):
        if getattr(self, "h", None):
            for p in self.problems:
                p.close()   # frees the views only
            lib.hpcgx_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_block(self, level, block):
        _check(lib.hpcgx_batch_set_block(self.h, level, block))

    def set_rows(self, level, rows):
        _check(lib.hpcgx_batch_set_rows(self.h, level, rows))

    def cg(self, max_iter=50, tol=0.0, mode=CG_F64, use_graph=True):
        n = len(self.problems)
        ni = np.zeros(n, np.int32)
        nr, nr0 = np.zeros(n), np.zeros(n)
        _check(lib.hpcgx_batch_cg(self.h, mode, max_iter, tol, int(use_graph), ni.ctypes.data,
                                  nr.ctypes.data, nr0.ctypes.data))
        return dict(niters=ni, normr=nr, normr0=nr0)

    BATCH_KINDS = ("spmv", "fwd_zero", "bwd_zero", "fwd", "bwd", "restrict", "prolong", "dot", "cg_p",
                   "cg_xr", "copy", "waxpby")

    def profile(self, iters=5, mode=CG_F64):
        """Per launch kind: (milliseconds per iteration, algorithmic bytes per
        iteration) from HIP events around every launch of `iters` eager batched
        iterations; plus milliseconds per iteration under "iteration"."""
        ms, by = np.zeros(12), np.zeros(12)
        it = C.c_double()
        _check(lib.hpcgx_batch_profile(self.h, mode, iters, ms.ctypes.data, by.ctypes.data, C.byref(it)))
        out = {k: (ms[i] / iters, by[i]) for i, k in enumerate(self.BATCH_KINDS)}
        out["iteration"] = (it.value, float(by.sum()))
        return out

    def history(self, k, max_len=4097):
        """normr after iterations 0..niters of problem k's last batched CG."""
        hist = np.zeros(max_len)
        n = C.c_int()
        _check(lib.hpcgx_batch_history(self.h, k, hist.ctypes.data, max_len, C.byref(n)))
        return hist[:n.value].copy()
