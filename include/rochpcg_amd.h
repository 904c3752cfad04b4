/*
 * rochpcg_amd.h -- C ABI of the MI355X-native HPCG solver hot path.
 *
 * A drop-in for the optimized CG/MG path of ROCm/rocHPCG on one GPU, single
 * rank.  Plain pointers and sizes only; every entry point returns 0 on success
 * and a non-zero status otherwise (hpcgx_last_error() holds the message) --
 * the library never calls exit(), unlike the reference's HIP_CHECK
 * (utils.hpp:63-76).
 *
 * Conventions shared with the reference's optimized path:
 *   - all device vectors are double precision and in multicolour row order
 *     (the order OptimizeProblem leaves them in, OptimizeProblem.cpp:69-160);
 *   - operators are asynchronous on the problem's stream; hpcgx_dot and
 *     hpcgx_fused_waxpby_dot return a host double and therefore synchronise,
 *     like ComputeDotProduct.cpp:216-217 and ComputeWAXPBY.cpp:218-219;
 *   - calls on one problem are not re-entrant (each problem owns its scratch).
 *
 * Each declaration names the reference interface it replaces (file:line of
 * ROCm/rocHPCG src/).
 */
#ifndef ROCHPCG_AMD_H
#define ROCHPCG_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hpcgx_problem hpcgx_problem;

/* precision stores built for a problem (bit mask) */
#define HPCGX_STORE_F64 1   /* f64 ELL values: the reference's arithmetic            */
#define HPCGX_STORE_MIXED 2 /* f16 ELL values + f32 MG vectors (mixed-precision CG)   */

/* rows per slice of the sliced ELL (hpcgx_level_arrays) */
#define HPCGX_ELL_SLICE 128

/* CG modes */
#define HPCGX_CG_F64 0      /* everything in f64                                      */
#define HPCGX_CG_MIXED 1    /* f64 outer CG, f32/f16 V-cycle, true residual every k   */

/* ---- setup ------------------------------------------------------------- */

/* GenerateGeometry + GenerateProblem + SetupHalo (single rank) +
 * GenerateCoarseProblem x (nlevels-1) + OptimizeProblem, all on the device:
 * GenerateProblem.cpp:375-489, GenerateCoarseProblem.cpp:98-184,
 * SetupHalo.cpp:313-317, OptimizeProblem.cpp:69-160 (JPLColoring,
 * PermuteColumns, ConvertToELL, PermuteRows, ExtractDiagonal).  The problem's
 * b = A*ones, x = 0, xexact = 1 are ready in multicolour order.
 * stream: a hipStream_t to run on, or NULL for a private non-blocking stream. */
int hpcgx_problem_create(int nx, int ny, int nz, int nlevels, int stores, void* stream,
                         hpcgx_problem** out);
/* DeleteMatrix / DeleteMGData / HIPDeleteCGData (SparseMatrix.hpp:264-330,
 * MGData.hpp:101-120, CGData.hpp:99-105) */
int hpcgx_problem_destroy(hpcgx_problem* p);

/* level geometry: info[0..7] = nx, ny, nz, rows, nonzeros, ncolors, coarse rows,
 * stores (SparseMatrix fields localNumberOfRows, totalNumberOfNonzeros,
 * nblocks; SparseMatrix.hpp:75-156) */
int hpcgx_level_info(const hpcgx_problem* p, int level, int64_t* info);
/* colour table: sizes[c], offsets[c] (c < ncolors, offsets has ncolors+1
 * entries) and the per-JPL-round counts the reference records in A.sizes
 * (MultiColoring.cpp:287-306), at most HPCGX_MAX_COLORS of them.  A level
 * holds up to HPCGX_MAX_COLORS colours, the reference's MAX_COLORS
 * (utils.hpp:48); a colouring that needs more fails problem creation. */
#define HPCGX_MAX_COLORS 128
int hpcgx_color_table(const hpcgx_problem* p, int level, int* sizes, int* offsets, int* round_sizes);
/* device pointers of a level's arrays (SparseMatrix::ell_col_ind, ell_val,
 * diag_idx, inv_diag, perm; MGData::d_f2cOperator); any output may be NULL.
 * The ELL is sliced: rows in slices of HPCGX_ELL_SLICE = 128, slice s holding
 * slot p of its rows at [(s*27 + p)*128, +128), i.e. entry p of row r at
 * ((r/128)*27 + p)*128 + r%128, over rows rounded up to a whole slice (the
 * reference's ELL is column-major, p*rows + r; SparseMatrix.cpp:271-407).  The
 * entries of each row are in the reference's ascending column order; the padding (-1, 0.0)
 * sits between the diagonal and the upper entries, which end at slot 26,
 * instead of at the row's tail (SparseMatrix.cpp:271-407 pads the tail).
 * ptrs must hold 13 entries: ptrs[0..12] = ell_col_ind, ell_val (f64 or
 * NULL), ell_val_f16 (or NULL), diag_idx, inv_diag, perm (natural ->
 * multicolour), colors (natural), f2c (coarse natural -> fine natural),
 * f2c_fine (coarse natural -> fine multicolour row), f2c_coarse (coarse
 * natural -> coarse multicolour row), row descriptors (int32: diagonal slot
 * in bits 0-7, first slot of an off-diagonal last-colour column in bits 8-15,
 * first upper slot in bits 16-23, 27 when absent; no reference counterpart), diagonal
 * values f64 (or NULL) and f16 (or NULL) (no reference counterpart) */
int hpcgx_level_arrays(const hpcgx_problem* p, int level, void** ptrs);
/* device pointers of the CG data and system: ptrs[0..7] = b, x, xexact, r,
 * z (f64, or NULL), p, Ap, z (f32 mixed-mode, or NULL)  (CGData.hpp:55-61) */
int hpcgx_vectors(const hpcgx_problem* p, void** ptrs);
/* CheckProblem (CheckProblem.cpp:78-179, run on every level by main.cpp:236-246)
 * on the device, on the optimized form of a level: counts[0..7] = errors in
 * the permutation (not a bijection), the colouring (not proper, colour table
 * or natural order inside a colour wrong), the ELL rows (not exactly the
 * row's in-range 27-point neighbours through perm, in strictly ascending
 * order, split-padded, with the closed-form count), the values (26 / -1 / 0
 * padding in every store, the diagonal arrays), the row descriptors and
 * diag_idx, and -- when check_vectors is set and level is 0 -- b = 26 -
 * (nnz_row - 1), x = 0, xexact = 1; then the stored nonzeros and the rows
 * checked.  Every row is checked, at any size.  Valid for the matrix as
 * generated (HIPReplaceMatrixDiagonal or TestCG's scaling change the values). */
int hpcgx_check_problem(hpcgx_problem* p, int level, int check_vectors, int64_t* counts);
/* refill b / xexact / x: kind 0 = the reference's b = A*ones, xexact = 1
 * (GenerateProblem.cpp:300); kind 1 = xexact uniform in [-1,1] from
 * splitmix64(seed), b = A*xexact summed in natural order; x = 0 */
int hpcgx_make_rhs(hpcgx_problem* p, int kind, uint64_t seed);

/* ---- operators (f64 device vectors, multicolour order) ------------------- */
/* A NULL vector (of a non-empty length) or result pointer is rejected with
 * -2 before anything is launched. */

/* ComputeSPMV(A, x, y)               ComputeSPMV.cpp:226-278 */
int hpcgx_spmv(hpcgx_problem* p, int level, const double* x, double* y);
/* ComputeSYMGS(A, r, x)              ComputeSYMGS.cpp:369-404 */
int hpcgx_symgs(hpcgx_problem* p, int level, const double* r, double* x);
/* ComputeSYMGSZeroGuess(A, r, x)     ComputeSYMGS.cpp:406-454 */
int hpcgx_symgs_zero(hpcgx_problem* p, int level, const double* r, double* x);
/* ComputeMG(A, r, x)                 ComputeMG.cpp:64-102 */
int hpcgx_mg(hpcgx_problem* p, int level, const double* r, double* x);
/* ComputeMG in the mixed store: r f64, x f32 (no reference counterpart; the
 * mixed-precision V-cycle of the mixed CG mode) */
int hpcgx_mg_mixed(hpcgx_problem* p, const double* r, float* x);
/* ComputeFusedSpMVRestriction(A, rf, xf): rc = restriction of rf - A xf into
 * the next level's rhs buffer (MGData::rc), ComputeRestriction.cpp:199-237;
 * rc_out (may be NULL) receives the device pointer of that buffer */
int hpcgx_restriction(hpcgx_problem* p, int level, const double* rf, const double* xf, double** rc_out);
/* The HPCG_REFERENCE build's unfused residual + restriction, as two calls:
 * ComputeSPMV(A, xf, Axf) with y == MGData::Axf computes A xf at the injected
 * rows only (ComputeSPMV.cpp:87-128, :238-256; the other rows of y are left
 * as they are), then ComputeRestriction(A, rf) forms rc = rf - Axf at those
 * rows into the next level's rhs buffer (ComputeRestriction.cpp:59-79,
 * :182-197); rc_out (may be NULL) receives that buffer's device pointer */
int hpcgx_spmv_injected(hpcgx_problem* p, int level, const double* xf, double* Axf);
int hpcgx_restriction_axf(hpcgx_problem* p, int level, const double* rf, const double* Axf, double** rc_out);
/* ComputeProlongation(Af, xf): xf += P xc with xc = the next level's
 * solution buffer (MGData::xc), ComputeProlongation.cpp:72-86; xc_out (may be
 * NULL) receives the device pointer of that buffer */
int hpcgx_prolongation(hpcgx_problem* p, int level, double* xf, double** xc_out);
/* ComputeDotProduct(n, x, y, result, ...)   ComputeDotProduct.cpp:189-238 */
int hpcgx_dot(hpcgx_problem* p, int64_t n, const double* x, const double* y, double* result);
/* hpcgx_dot without a problem handle and without any device allocation.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The same kernels as hpcgx_dot, with the block partials and the
 * result in a fixed 33 KB device array of the library (4096 partials; the
 * reduction never uses more, whatever n); a call that would need more is
 * refused.  Runs on `stream` (NULL: the null stream) of the current device
 * and synchronises it; *result is a host double.  Every call on one GPU
 * shares that array: calls that overlap on one GPU race. */
int hpcgx_dot_noalloc(int64_t n, const double* x, const double* y, double* result, void* stream);
/* ComputeResidual(n, v1, v2, residual): max_i |v1_i - v2_i|
 * ComputeResidual.cpp:121-152 (synchronises, like the reference) */
int hpcgx_residual(hpcgx_problem* p, int64_t n, const double* v1, const double* v2, double* result);
/* ComputeWAXPBY(n, alpha, x, beta, y, w)    ComputeWAXPBY.cpp:111-135 */
int hpcgx_waxpby(hpcgx_problem* p, int64_t n, double alpha, const double* x, double beta,
                 const double* y, double* w);
/* ComputeFusedWAXPBYDot(n, alpha, x, y, result): y += alpha x; result = y'y
 * ComputeWAXPBY.cpp:198-240 */
int hpcgx_fused_waxpby_dot(hpcgx_problem* p, int64_t n, double alpha, const double* x, double* y,
                           double* result);
/* HIPCopyMatrixDiagonal / HIPReplaceMatrixDiagonal (SparseMatrix.cpp:102-164),
 * f64 store only */
int hpcgx_copy_diagonal(hpcgx_problem* p, int level, double* diag);
int hpcgx_replace_diagonal(hpcgx_problem* p, int level, const double* diag);
/* TestCG's diagonal/rhs exaggeration (TestCG.cpp:67-95): d[i], b[i] *= 1e6,
 * (i+2)*1e6 for i < 9 */
int hpcgx_testcg_scale(hpcgx_problem* p, double* diag, double* b);

/* ---- CG --------------------------------------------------------------- */

/* CG(A, data, b, x, max_iter, tolerance, niters, normr, normr0, times,
 *    doPreconditioning)  CG.cpp:104-200.
 * b, x: device vectors (NULL = the problem's own b / x).  history (host,
 * may be NULL) receives normr after iterations 0..niters (hist_len entries at
 * most).  times (host, may be NULL): times[0] += wall seconds.
 * use_graph: replay each iteration as a captured hipGraph. */
int hpcgx_cg(hpcgx_problem* p, int mode, const double* b, double* x, int max_iter, double tolerance,
             int precond, int use_graph, int* niters, double* normr, double* normr0,
             double* history, int hist_len, double* times);
/* the same loop in pieces, for timing: preamble; `count` iterations enqueued
 * (no synchronisation); results.  Each iteration's x += alpha p is applied by
 * the next one, so x is complete after hpcgx_cg_end (which applies the last)
 * and lags one update in between. */
int hpcgx_cg_begin(hpcgx_problem* p, int mode, const double* b, double* x, int max_iter,
                   double tolerance, int precond);
int hpcgx_cg_iterate(hpcgx_problem* p, int count, int use_graph);
/* `count` iterations with HIP events around each phase; phase_ms[0..4] =
 * SpMV(+p'Ap), MG, r'z, vector updates, true-residual recomputation */
int hpcgx_cg_iterate_profiled(hpcgx_problem* p, int count, double* phase_ms);
int hpcgx_cg_end(hpcgx_problem* p, int* niters, double* normr, double* normr0, double* history,
                 int hist_len);
/* mixed mode: recompute the true residual every k iterations (0 = never, the
 * default: the outer CG's operator is exact, 26 and -1 being exact in f16, so
 * its recursive residual tracks the true one as in the FP64 CG) */
int hpcgx_set_residual_replacement(hpcgx_problem* p, int every);

/* HPCG flop count of one CG iteration (ReportResults.cpp:73-89) */
double hpcgx_flops_per_iteration(const hpcgx_problem* p);
/* algorithmic HBM bytes of one V-cycle and of one CG iteration in `mode`
 * (HPCGX_CG_F64 / HPCGX_CG_MIXED) as this library's kernels run them: each
 * ELL slot a kernel reads counted once, neighbour gathers from cache, own-row
 * vectors once (no reference counterpart; the reference's traffic model is
 * ReportResults.cpp:107-138) */
int hpcgx_iteration_bytes(const hpcgx_problem* p, int mode, double* vcycle_bytes, double* iteration_bytes);

/* ---- the reference's natural-order path on the device -------------------- */
/* Natural row order (GenerateProblem's numbering), the unoptimized operators
 * of the reference with its operation order and rounding (no fused
 * multiply-add); the Gauss-Seidel sweeps are the reference's sequential
 * sweeps, run as wavefronts ix + 2 iy + 4 iz (every lower neighbour of a row
 * lies in an earlier wavefront).  The benchmark driver takes its reference
 * tolerance from hpcgx_cg_ref, as main.cpp:334-344 takes it from CG_ref.
 * All vectors are f64 device vectors in natural order. */

/* ComputeSPMV_ref(A, x, y)           ComputeSPMV_ref.cpp:47-72 */
int hpcgx_spmv_ref(hpcgx_problem* p, int level, const double* x, double* y);
/* ComputeSYMGS_ref(A, r, x)          ComputeSYMGS_ref.cpp:54-103 */
int hpcgx_symgs_ref(hpcgx_problem* p, int level, const double* r, double* x);
/* ComputeMG_ref(A, r, x)             ComputeMG_ref.cpp:39-63 */
int hpcgx_mg_ref(hpcgx_problem* p, int level, const double* r, double* x);
/* CG_ref(A, data, b, x, max_iter, tolerance, niters, normr, normr0, times,
 *        doPreconditioning)          CG_ref.cpp:86-174
 * b NULL = the problem's b taken back to natural order; x NULL = a work
 * vector of the path, zeroed first (else x is the initial guess and receives
 * the solution).  history/hist_len/times as in hpcgx_cg. */
int hpcgx_cg_ref(hpcgx_problem* p, const double* b, double* x, int max_iter, double tolerance, int precond,
                 int* niters, double* normr, double* normr0, double* history, int hist_len, double* times);

/* ---- batched solves (no reference counterpart) --------------------------- */

/* Many independent problems of uneven sizes, each with its own hierarchy
 * (dims: nprob triples nx, ny, nz).  Every operator of the V-cycle and of the
 * CG runs as one launch per (operator, level, colour) across all problems,
 * with per-problem reductions, scalars and convergence; one batched CG
 * iteration is one hipGraph. */
typedef struct hpcgx_batch hpcgx_batch;
int hpcgx_batch_create(int nprob, const int* dims, int nlevels, int stores, void* stream, hpcgx_batch** out);
int hpcgx_batch_destroy(hpcgx_batch* b);
/* a non-owning handle on problem k (its vectors, rhs, per-problem operators);
 * release it with hpcgx_problem_destroy, which then frees only the handle */
int hpcgx_batch_problem(hpcgx_batch* b, int k, hpcgx_problem** out);
/* threads per block of every batched kernel on `level` (64, 128 or 256) */
int hpcgx_batch_set_block(hpcgx_batch* b, int level, int block);
/* rows per thread (1 or 2) of a level's batched sweep launches; a launch whose
 * segments are not all even in offset, size and row count runs with 1 (no reference
 * counterpart; the autotuner's knob) */
int hpcgx_batch_set_rows(hpcgx_batch* b, int level, int rows);
/* CG on every problem from its own x (b, x of each problem); niters, normr,
 * normr0: arrays of nprob entries (may be NULL) */
int hpcgx_batch_cg(hpcgx_batch* b, int mode, int max_iter, double tolerance, int use_graph, int* niters,
                   double* normr, double* normr0);
/* `iters` batched CG iterations from each problem's x, launched one by one
 * with HIP events around every launch: kind_ms[k] (12 entries) = summed
 * milliseconds of the launches of kind k over the iterations -- 0 SpMV,
 * 1/2 zero-guess forward/backward sweeps, 3/4 forward/backward sweeps,
 * 5 restriction, 6 prolongation, 7 dot, 8 p update, 9 x/r update, 10 copy,
 * 11 waxpby -- kind_bytes[k] = their algorithmic HBM bytes per iteration,
 * *iter_ms = milliseconds per iteration (no reference counterpart) */
int hpcgx_batch_profile(hpcgx_batch* b, int mode, int iters, double* kind_ms, double* kind_bytes, double* iter_ms);
/* residual history of problem k's last batched CG: normr after iterations
 * 0..niters into history (at most hist_len entries); *count = entries written */
int hpcgx_batch_history(hpcgx_batch* b, int k, double* history, int hist_len, int* count);

/* ---- timing and memory helpers ----------------------------------------- */

/* average device time (ms) of `reps` back-to-back runs of one operator on
 * `level`, measured with HIP events on the problem's stream, with the
 * operands the V-cycle uses there (level 0: the CG residual r and z).
 * op 0 = SpMV, 1 = SymGS, 2 = zero-guess SymGS, 3 = MG from `level`,
 * 4/5 = zero-guess forward/backward sweeps, 6/7 = forward/backward sweeps
 * (in the form the V-cycle runs them on that level), 8 = fused residual +
 * restriction, 9 = one read of the level's ELL column and value arrays with
 * 16-byte loads and no gathers (the read ceiling of the matrix stream that
 * the SpMV reads: 27 * (4 + value bytes) per row); prec 0 = f64 store,
 * 1 = mixed store (SpMV: f16 values with f64 vectors).  SpMV on level 0 is
 * timed as the CG runs it: with the fused p'Ap partials and their fold. */
int hpcgx_time_operator(hpcgx_problem* p, int op, int level, int prec, int reps, double* avg_ms);
/* launch shape of an operator on a level (the autotuner's knobs; no
 * reference counterpart -- the reference hard-codes 1024/256/128-thread
 * blocks, ComputeSPMV.cpp:262, ComputeSYMGS.cpp:395): op 0 = SpMV,
 * 1 = zero-guess forward sweep, 2 = zero-guess backward sweep, 3 = forward
 * sweep, 4 = backward sweep, 5 = restriction, 6-9 = ops 1-4 on their
 * narrow colours (slot window of at most 16 slots; ops 1-4 then apply to
 * the wide colours); block in {64,128,256} threads, rows in {1,2}
 * consecutive rows per thread (restriction: 1).  level -1 / op -1 apply to
 * all; prec 0 = f64 kernels, 1 = mixed kernels.  Colour blocks not
 * divisible by `rows` run with 1. */
int hpcgx_set_shape(hpcgx_problem* p, int prec, int level, int op, int block, int rows);
int hpcgx_get_shape(const hpcgx_problem* p, int prec, int level, int op, int* block, int* rows);
/* Block order of the SymGS colour launches (no reference counterpart; a
 * tuning knob like the launch shape).  columns 0 = choose (planes of the
 * half-resolution colour grid too large for an XCD's L2 are swept column by
 * column), 1 = row order, k > 1 = k columns per plane where the colour's tiles
 * divide so.  Results do not depend on it.  level -1 applies to all. */
int hpcgx_set_sweep_columns(hpcgx_problem* p, int level, int columns);
int hpcgx_synchronize(hpcgx_problem* p);
int hpcgx_malloc(void** ptr, size_t bytes);
int hpcgx_free(void* ptr);
int hpcgx_memcpy_h2d(void* dst, const void* src, size_t bytes);
int hpcgx_memcpy_d2h(void* dst, const void* src, size_t bytes);
int hpcgx_memcpy_d2d(void* dst, const void* src, size_t bytes);
int hpcgx_device_count(int* count);
const char* hpcgx_last_error(void);
const char* hpcgx_version(void);

#ifdef __cplusplus
}
#endif
#endif /* ROCHPCG_AMD_H */
