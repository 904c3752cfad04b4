// ops.hpp -- host launchers of the hot-path kernels (ops.hip).
#pragma once
#include "common.hpp"

namespace hpcgx
{

// What the last block of a reduction does with the folded sum.
enum RedOp : int
{
    OP_NONE = 0,
    OP_STORE = 1,   // *out = sum
    OP_NORMR0 = 2,  // normr = normr0 = sqrt(sum); iteration 0 of the history
    OP_RTZ = 3,     // oldrtz = rtz; rtz = sum; beta = rtz/oldrtz (after the first iteration)
    OP_PAP = 4,     // pAp = sum; alpha = rtz/pAp
    OP_NORMR = 5,   // normr = sqrt(sum); iter += 1; history; convergence flag
    OP_NORMR_SET = 6 // normr = sqrt(sum) of a recomputed true residual (same iteration)
};

struct RedArgs
{
    double* partial = nullptr;      // one slot per block of the producing kernel
    int op = OP_NONE;
    CGScalars* sc = nullptr;
    double* hist = nullptr;
    int hist_cap = -1;
    double* out = nullptr;
    int64_t cap = RED_BLOCKS;      // partial slots available
};

bool shape_supported(int rows, int block);

// one-block fold of `nparts` partials written by the previous launch on `s`
int launch_fold(int64_t nparts, const RedArgs& ra, hipStream_t s, const int* gate);

// `gate` (device int, may be null): kernels return at once while *gate != 0.
int launch_spmv(const Level& L, Prec prec, const void* x, void* y, hipStream_t s, const int* gate);
int launch_spmv_f64x(const Level& L, Prec prec, const double* x, double* y, const RedArgs* ra, hipStream_t s,
                     const int* gate);
// slots a sweep loads per row for a window of w slots (its compile-time
// bucket), and for a (lower, last-colour) pair of the post-smoother's forward
// sweep
int window_slots(int w);
void window_pair_slots(int wl, int wt, int* a, int* b);
// one sweep of a level: 0 zero-guess fwd, 1 zero-guess bwd, 2 fwd keeping t, 3 bwd from t,
// 4 = 1 also keeping the upper sums outside the last colour in t, 5 = 2 reusing them
int launch_sweep(const Level& L, Prec prec, int part, const void* r, bool r_f64, void* x, void* t, hipStream_t s,
                 const int* gate);
// u (may be null): also keep the upper sums outside the last colour, for a
// post-smoother launch_symgs(..., use_u = true) after a coarse-grid
// correction that changes only last-colour rows (Level::u_ok)
int launch_symgs_zero(const Level& L, Prec prec, const void* r, bool r_f64, void* x, hipStream_t s, const int* gate,
                      void* u = nullptr);
int launch_symgs(const Level& L, Prec prec, const void* r, bool r_f64, void* x, void* t, hipStream_t s, const int* gate,
                 bool use_u = false);
int launch_restrict(const Level& F, Prec prec, const void* r, bool r_f64, const void* x, void* rc, hipStream_t s,
                    const int* gate);
int launch_prolong(const Level& F, Prec prec, const void* xc, void* x, hipStream_t s, const int* gate);
int launch_spmv_injected(const Level& F, const double* x, double* y, hipStream_t s);
int launch_restrict_diff(const Level& F, const double* rf, const double* Axf, double* rc, hipStream_t s);
// CheckProblem on the device (check.hip): counts[CHECK_*] of a level's
// optimized form -- perm, colouring, ELL rows, values, descriptors and (when
// b/x/xexact are given) the default right-hand side; counts[CHECK_NNZ] the
// stored nonzeros, counts[CHECK_ROWS] the rows checked
enum CheckKind : int
{
    CHECK_PERM = 0,
    CHECK_COLOR = 1,
    CHECK_ELL = 2,
    CHECK_VALUE = 3,
    CHECK_DESC = 4,
    CHECK_VECTORS = 5,
    CHECK_NNZ = 6,
    CHECK_ROWS = 7,
    CHECK_KINDS = 8
};
int check_problem(const Level& L, const double* b, const double* x, const double* xexact, int64_t* counts,
                  hipStream_t s);
// one read of a level's column and value arrays (16-byte non-temporal loads,
// no gathers): the read ceiling of the matrix stream; sink holds >= min(4096,
// sink_cap) doubles and is not normally written
int launch_read_stream(const Level& L, Prec prec, double* sink, int64_t sink_cap, hipStream_t s);
// blocks (= partial slots) of a streaming vector reduction over n elements:
// at most RED_BLOCKS whatever n
int64_t reduction_blocks(int64_t n);
int launch_dot(int64_t n, const void* x, bool x_f64, const void* y, bool y_f64, const RedArgs& ra, hipStream_t s,
               const int* gate);
// the dot product over a file-scope device scratch array (noalloc.hip): no
// device allocation; *result (host) is ready on return
int dot_noalloc(int64_t n, const double* x, const double* y, double* result, hipStream_t s);
// *out = max_i |x_i - y_i| (two launches; partial holds >= min(blocks, cap) slots)
int launch_absdiff_max(int64_t n, const double* x, const double* y, double* partial, int64_t cap, double* out,
                       hipStream_t s);
int launch_waxpby(int64_t n, double alpha, const double* x, double beta, const double* y, double* w, hipStream_t s,
                  const int* gate);
int launch_axpy_dot(int64_t n, double alpha, const double* x, double* y, const RedArgs& ra, hipStream_t s,
                    const int* gate);
// p = z + beta p, applying first the x += alpha p that the previous
// launch_cg_xr left pending (CGScalars::xpend)
int launch_cg_p(int64_t n, const void* z, bool z_f64, double* p, double* x, const CGScalars* sc, hipStream_t s,
                const int* gate);
// r -= alpha Ap and ||r||; x += alpha p stays pending
int launch_cg_xr(int64_t n, const double* Ap, double* r, const RedArgs& ra, hipStream_t s, const int* gate);
// apply a pending x += alpha p (before x is read: true residual, end of CG)
int launch_cg_xflush(int64_t n, const double* p, double* x, const CGScalars* sc, hipStream_t s, const int* gate);
int launch_fill(int64_t n, void* p, bool f64, double v, hipStream_t s, const int* gate);
int launch_copy(int64_t n, const void* src, bool src_f64, void* dst, bool dst_f64, hipStream_t s, const int* gate);
int launch_get_diag(const Level& L, double* d, hipStream_t s);
int launch_set_diag(Level& L, const double* d, hipStream_t s);
int launch_testcg_scale(int64_t m, double* d, double* b, hipStream_t s);

}  // namespace hpcgx
