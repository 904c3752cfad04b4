// capi.cpp -- the extern "C" surface (include/rochpcg_amd.h) over the solver.
#include "../../include/rochpcg_amd.h"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <string>

#include "batch.hpp"
#include "ops.hpp"
#include "refpath.hpp"
#include "setup.hpp"
#include "solver.hpp"

namespace hpcgx
{
static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }
const char* last_error() { return g_err.c_str(); }
}  // namespace hpcgx

using namespace hpcgx;

struct hpcgx_problem
{
    Problem* P;
    bool owned;   // false for a view of one problem of a batch
};

struct hpcgx_batch
{
    Batch* B;
};

#define CHECK_P(p)                                                                                                     \
    do                                                                                                                 \
    {                                                                                                                  \
        if(!(p) || !(p)->P)                                                                                            \
        {                                                                                                              \
            set_error("null problem handle");                                                                          \
            return -2;                                                                                                 \
        }                                                                                                              \
    } while(0)

#define CHECK_LEVEL(p, l)                                                                                              \
    do                                                                                                                 \
    {                                                                                                                  \
        if((l) < 0 || (l) >= (p)->P->nlevels)                                                                          \
        {                                                                                                              \
            set_error("level out of range");                                                                           \
            return -2;                                                                                                 \
        }                                                                                                              \
    } while(0)

// Null device or host pointers are turned into an error here: a kernel given
// one would fault the device instead of returning.
#define CHECK_PTRS(...)                                                                                                \
    do                                                                                                                 \
    {                                                                                                                  \
        const void* ptrs_[] = {__VA_ARGS__};                                                                           \
        for(const void* q_ : ptrs_)                                                                                    \
            if(!q_)                                                                                                    \
            {                                                                                                          \
                set_error("null pointer argument");                                                                    \
                return -2;                                                                                             \
            }                                                                                                          \
    } while(0)

#define NEED_F64(p)                                                                                                    \
    do                                                                                                                 \
    {                                                                                                                  \
        if(!(p)->P->has_f64)                                                                                           \
        {                                                                                                              \
            set_error("operation needs the f64 store (HPCGX_STORE_F64)");                                              \
            return -2;                                                                                                 \
        }                                                                                                              \
    } while(0)

extern "C" {

int hpcgx_problem_create(int nx, int ny, int nz, int nlevels, int stores, void* stream, hpcgx_problem** out)
{
    if(!out)
    {
        set_error("null output pointer");
        return -2;
    }
    *out = nullptr;
    Problem* P = nullptr;
    int rc = problem_create(P, nx, ny, nz, nlevels, (stores & HPCGX_STORE_F64) != 0, (stores & HPCGX_STORE_MIXED) != 0,
                            (hipStream_t)stream);
    if(rc) return rc;
    *out = new hpcgx_problem{P, true};
    return 0;
}

int hpcgx_problem_destroy(hpcgx_problem* p)
{
    if(!p) return 0;
    if(p->owned) problem_destroy(p->P);
    delete p;
    return 0;
}

int hpcgx_level_info(const hpcgx_problem* p, int level, int64_t* info)
{
    CHECK_P(p);
    CHECK_LEVEL(p, level);
    CHECK_PTRS(info);
    const Level& L = p->P->L[level];
    info[0] = L.nx;
    info[1] = L.ny;
    info[2] = L.nz;
    info[3] = L.m;
    info[4] = L.nnz;
    info[5] = L.ncolors;
    info[6] = L.mc;
    info[7] = (L.ell_val ? HPCGX_STORE_F64 : 0) | (L.ell_h ? HPCGX_STORE_MIXED : 0);
    return 0;
}

int hpcgx_color_table(const hpcgx_problem* p, int level, int* sizes, int* offsets, int* round_sizes)
{
    CHECK_P(p);
    CHECK_LEVEL(p, level);
    const Level& L = p->P->L[level];
    for(int c = 0; c < L.ncolors; ++c)
    {
        if(sizes) sizes[c] = L.sizes[c];
        if(offsets) offsets[c] = L.offsets[c];
    }
    if(offsets) offsets[L.ncolors] = L.offsets[L.ncolors];
    if(round_sizes)
    {
        int nr = 0;
        while(nr < MAX_COLORS && (nr < L.ncolors || L.round_sizes[nr] != 0)) ++nr;
        for(int k = 0; k < nr; ++k) round_sizes[k] = L.round_sizes[k];
    }
    return 0;
}

int hpcgx_level_arrays(const hpcgx_problem* p, int level, void** ptrs)
{
    CHECK_P(p);
    CHECK_LEVEL(p, level);
    CHECK_PTRS(ptrs);
    const Level& L = p->P->L[level];
    void* v[] = {L.ell_col, L.ell_val, L.ell_h, L.diag_idx, L.inv_diag, L.perm, L.colors, L.f2c_nat, L.f2c_fine,
                 L.f2c_coarse, L.desc, L.diag64, L.diag_h};
    for(int i = 0; i < 13; ++i) ptrs[i] = v[i];
    return 0;
}

int hpcgx_vectors(const hpcgx_problem* p, void** ptrs)
{
    CHECK_P(p);
    CHECK_PTRS(ptrs);
    Problem* P = p->P;
    void* v[] = {P->b, P->x, P->xexact, P->r, P->z64, P->p, P->Ap, P->z32};
    for(int i = 0; i < 8; ++i) ptrs[i] = v[i];
    return 0;
}

int hpcgx_check_problem(hpcgx_problem* p, int level, int check_vectors, int64_t* counts)
{
    CHECK_P(p);
    CHECK_LEVEL(p, level);
    CHECK_PTRS(counts);
    Problem* P = p->P;
    const bool vec = check_vectors != 0 && level == 0;
    return check_problem(P->L[level], vec ? P->b : nullptr, vec ? P->x : nullptr, vec ? P->xexact : nullptr, counts,
                         P->s);
}

int hpcgx_make_rhs(hpcgx_problem* p, int kind, uint64_t seed)
{
    CHECK_P(p);
    if(kind != 0 && kind != 1)
    {
        set_error("rhs kind must be 0 or 1");
        return -2;
    }
    HPCGX_TRY(make_rhs(p->P->L[0], kind, seed, p->P->b, p->P->xexact, p->P->x, p->P->s));
    return 0;
}

// ---------------------------------------------------------------------------
int hpcgx_spmv(hpcgx_problem* p, int level, const double* x, double* y)
{
    CHECK_P(p);
    CHECK_LEVEL(p, level);
    CHECK_PTRS(x, y);
    const Level& L = p->P->L[level];
    // f64 vectors over the f64 store, or over the exact f16 store when that is all there is
    return launch_spmv_f64x(L, L.ell_val ? PREC_F64 : PREC_MIXED, x, y, nullptr, p->P->s, nullptr);
}

int hpcgx_symgs(hpcgx_problem* p, int level, const double* r, double* x)
{
    CHECK_P(p);
    CHECK_LEVEL(p, level);
    CHECK_PTRS(r, x);
    NEED_F64(p);
    return launch_symgs(p->P->L[level], PREC_F64, r, true, x, p->P->v64[level].t, p->P->s, nullptr);
}

int hpcgx_symgs_zero(hpcgx_problem* p, int level, const double* r, double* x)
{
    CHECK_P(p);
    CHECK_LEVEL(p, level);
    CHECK_PTRS(r, x);
    NEED_F64(p);
    return launch_symgs_zero(p->P->L[level], PREC_F64, r, true, x, p->P->s, nullptr);
}

int hpcgx_mg(hpcgx_problem* p, int level, const double* r, double* x)
{
    CHECK_P(p);
    CHECK_LEVEL(p, level);
    CHECK_PTRS(r, x);
    NEED_F64(p);
    return mg(*p->P, level, PREC_F64, r, true, x, nullptr);
}

int hpcgx_mg_mixed(hpcgx_problem* p, const double* r, float* x)
{
    CHECK_P(p);
    CHECK_PTRS(r, x);
    if(!p->P->has_mixed)
    {
        set_error("operation needs the mixed store (HPCGX_STORE_MIXED)");
        return -2;
    }
    return mg(*p->P, 0, PREC_MIXED, r, true, x, nullptr);
}

int hpcgx_restriction(hpcgx_problem* p, int level, const double* rf, const double* xf, double** rc_out)
{
    CHECK_P(p);
    CHECK_LEVEL(p, level);
    CHECK_PTRS(rf, xf);
    NEED_F64(p);
    if(level + 1 >= p->P->nlevels)
    {
        set_error("the coarsest level has no restriction");
        return -2;
    }
    void* rc = p->P->v64[level + 1].r;
    if(rc_out) *rc_out = (double*)rc;
    return launch_restrict(p->P->L[level], PREC_F64, rf, true, xf, rc, p->P->s, nullptr);
}

int hpcgx_spmv_injected(hpcgx_problem* p, int level, const double* xf, double* Axf)
{
    CHECK_P(p);
    CHECK_LEVEL(p, level);
    CHECK_PTRS(xf, Axf);
    NEED_F64(p);
    if(level + 1 >= p->P->nlevels)
    {
        set_error("the coarsest level has no injected rows");
        return -2;
    }
    return launch_spmv_injected(p->P->L[level], xf, Axf, p->P->s);
}

int hpcgx_restriction_axf(hpcgx_problem* p, int level, const double* rf, const double* Axf, double** rc_out)
{
    CHECK_P(p);
    CHECK_LEVEL(p, level);
    CHECK_PTRS(rf, Axf);
    NEED_F64(p);
    if(level + 1 >= p->P->nlevels)
    {
        set_error("the coarsest level has no restriction");
        return -2;
    }
    double* rc = (double*)p->P->v64[level + 1].r;
    if(rc_out) *rc_out = rc;
    return launch_restrict_diff(p->P->L[level], rf, Axf, rc, p->P->s);
}

int hpcgx_prolongation(hpcgx_problem* p, int level, double* xf, double** xc_out)
{
    CHECK_P(p);
    CHECK_LEVEL(p, level);
    CHECK_PTRS(xf);
    NEED_F64(p);
    if(level + 1 >= p->P->nlevels)
    {
        set_error("the coarsest level has no prolongation");
        return -2;
    }
    void* xc = p->P->v64[level + 1].x;
    if(xc_out) *xc_out = (double*)xc;
    return launch_prolong(p->P->L[level], PREC_F64, xc, xf, p->P->s, nullptr);
}

static RedArgs store_red(Problem* P)
{
    RedArgs ra;
    ra.partial = P->partial;
    ra.op = OP_STORE;
    ra.out = P->dout;
    ra.cap = P->partial_cap;
    return ra;
}

int hpcgx_dot(hpcgx_problem* p, int64_t n, const double* x, const double* y, double* result)
{
    CHECK_P(p);
    CHECK_PTRS(result);
    if(n > 0) CHECK_PTRS(x, y);
    Problem* P = p->P;
    HPCGX_TRY(launch_dot(n, x, true, y, true, store_red(P), P->s, nullptr));
    HPCGX_HIP(hipMemcpyAsync(P->h_out, P->dout, sizeof(double), hipMemcpyDeviceToHost, P->s));
    HPCGX_HIP(hipStreamSynchronize(P->s));
    *result = P->h_out[0];
    return 0;
}

int hpcgx_dot_noalloc(int64_t n, const double* x, const double* y, double* result, void* stream)
{
    CHECK_PTRS(result);
    if(n > 0) CHECK_PTRS(x, y);
    return dot_noalloc(n, x, y, result, (hipStream_t)stream);
}

int hpcgx_residual(hpcgx_problem* p, int64_t n, const double* v1, const double* v2, double* result)
{
    CHECK_P(p);
    CHECK_PTRS(result);
    if(n > 0) CHECK_PTRS(v1, v2);
    Problem* P = p->P;
    HPCGX_TRY(launch_absdiff_max(n, v1, v2, P->partial, P->partial_cap, P->dout, P->s));
    HPCGX_HIP(hipMemcpyAsync(P->h_out, P->dout, sizeof(double), hipMemcpyDeviceToHost, P->s));
    HPCGX_HIP(hipStreamSynchronize(P->s));
    *result = P->h_out[0];
    return 0;
}

int hpcgx_waxpby(hpcgx_problem* p, int64_t n, double alpha, const double* x, double beta, const double* y, double* w)
{
    CHECK_P(p);
    if(n > 0) CHECK_PTRS(x, y, w);
    return launch_waxpby(n, alpha, x, beta, y, w, p->P->s, nullptr);
}

int hpcgx_fused_waxpby_dot(hpcgx_problem* p, int64_t n, double alpha, const double* x, double* y, double* result)
{
    CHECK_P(p);
    CHECK_PTRS(result);
    if(n > 0) CHECK_PTRS(x, y);
    Problem* P = p->P;
    HPCGX_TRY(launch_axpy_dot(n, alpha, x, y, store_red(P), P->s, nullptr));
    HPCGX_HIP(hipMemcpyAsync(P->h_out, P->dout, sizeof(double), hipMemcpyDeviceToHost, P->s));
    HPCGX_HIP(hipStreamSynchronize(P->s));
    *result = P->h_out[0];
    return 0;
}

int hpcgx_copy_diagonal(hpcgx_problem* p, int level, double* diag)
{
    CHECK_P(p);
    CHECK_LEVEL(p, level);
    CHECK_PTRS(diag);
    return launch_get_diag(p->P->L[level], diag, p->P->s);
}

int hpcgx_replace_diagonal(hpcgx_problem* p, int level, const double* diag)
{
    CHECK_P(p);
    CHECK_LEVEL(p, level);
    CHECK_PTRS(diag);
    return launch_set_diag(p->P->L[level], diag, p->P->s);
}

int hpcgx_testcg_scale(hpcgx_problem* p, double* diag, double* b)
{
    CHECK_P(p);
    CHECK_PTRS(diag, b);
    return launch_testcg_scale(p->P->L[0].m, diag, b, p->P->s);
}

// ---------------------------------------------------------------------------
int hpcgx_cg(hpcgx_problem* p, int mode, const double* b, double* x, int max_iter, double tolerance, int precond,
             int use_graph, int* niters, double* normr, double* normr0, double* history, int hist_len, double* times)
{
    CHECK_P(p);
    const auto t0 = std::chrono::steady_clock::now();
    int rc = cg_solve(*p->P, mode, precond != 0, b, x, max_iter, tolerance, use_graph != 0, niters, normr, normr0,
                      history, hist_len);
    if(times) times[0] += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return rc;
}

int hpcgx_cg_begin(hpcgx_problem* p, int mode, const double* b, double* x, int max_iter, double tolerance, int precond)
{
    CHECK_P(p);
    return cg_begin(*p->P, mode, precond != 0, b, x, max_iter, tolerance);
}

int hpcgx_cg_iterate(hpcgx_problem* p, int count, int use_graph)
{
    CHECK_P(p);
    return cg_iterations(*p->P, count, use_graph != 0);
}

int hpcgx_cg_iterate_profiled(hpcgx_problem* p, int count, double* phase_ms)
{
    CHECK_P(p);
    CHECK_PTRS(phase_ms);
    return cg_iterations_profiled(*p->P, count, phase_ms);
}

int hpcgx_cg_end(hpcgx_problem* p, int* niters, double* normr, double* normr0, double* history, int hist_len)
{
    CHECK_P(p);
    return cg_end(*p->P, niters, normr, normr0, history, hist_len);
}

// ---------------------------------------------------------------------------
int hpcgx_spmv_ref(hpcgx_problem* p, int level, const double* x, double* y)
{
    CHECK_P(p);
    CHECK_LEVEL(p, level);
    CHECK_PTRS(x, y);
    return ref_spmv(*p->P, level, x, y);
}

int hpcgx_symgs_ref(hpcgx_problem* p, int level, const double* r, double* x)
{
    CHECK_P(p);
    CHECK_LEVEL(p, level);
    CHECK_PTRS(r, x);
    return ref_symgs(*p->P, level, r, x);
}

int hpcgx_mg_ref(hpcgx_problem* p, int level, const double* r, double* x)
{
    CHECK_P(p);
    CHECK_LEVEL(p, level);
    CHECK_PTRS(r, x);
    return ref_mg(*p->P, level, r, x);
}

int hpcgx_cg_ref(hpcgx_problem* p, const double* b, double* x, int max_iter, double tolerance, int precond,
                 int* niters, double* normr, double* normr0, double* history, int hist_len, double* times)
{
    CHECK_P(p);
    const auto t0 = std::chrono::steady_clock::now();
    int rc = ref_cg(*p->P, b, x, max_iter, tolerance, precond != 0, niters, normr, normr0, history, hist_len);
    if(times) times[0] += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return rc;
}

int hpcgx_set_residual_replacement(hpcgx_problem* p, int every)
{
    CHECK_P(p);
    if(every < 0)
    {
        set_error("replacement interval must be >= 0");
        return -2;
    }
    p->P->replace_every = every;
    return 0;
}

double hpcgx_flops_per_iteration(const hpcgx_problem* p)
{
    if(!p || !p->P) return 0.0;
    return hpcg_flops_per_iter(*p->P);
}

int hpcgx_iteration_bytes(const hpcgx_problem* p, int mode, double* vcycle_bytes, double* iteration_bytes)
{
    CHECK_P(p);
    if(mode != CG_F64 && mode != CG_MIXED)
    {
        set_error("invalid CG mode");
        return -2;
    }
    return hpcgx::iteration_bytes(*p->P, mode, vcycle_bytes, iteration_bytes);
}

// ---------------------------------------------------------------------------
int hpcgx_time_operator(hpcgx_problem* p, int op, int level, int prec, int reps, double* avg_ms)
{
    CHECK_P(p);
    CHECK_LEVEL(p, level);
    CHECK_PTRS(avg_ms);
    Problem* P = p->P;
    const Level& L = P->L[level];
    const bool mixed = prec != 0;
    if((mixed && !P->has_mixed) || (!mixed && !P->has_f64) || op < 0 || op > 9 || reps < 1)
    {
        set_error("operator not available for this problem");
        return -2;
    }
    if(op == 8 && level + 1 >= P->nlevels)
    {
        set_error("the coarsest level has no restriction");
        return -2;
    }
    // operands as the V-cycle uses them: level 0 takes the CG residual (f64)
    // and writes z; coarser levels use their own r / x buffers
    const Prec pr = mixed ? PREC_MIXED : PREC_F64;
    LevelVecs* V = mixed ? P->v32 : P->v64;
    const void* r = (level == 0) ? (const void*)P->r : V[level].r;
    void* x = (level == 0) ? (mixed ? (void*)P->z32 : (void*)P->z64) : V[level].x;
    const bool r_f64 = (level == 0) || !mixed;
    auto run = [&]() -> int {
        switch(op)
        {
        case 0:
        {
            // level 0 runs as the CG runs it: with the fused p'Ap partials and their fold
            RedArgs ra = store_red(P);
            const RedArgs* dot = (level == 0) ? &ra : nullptr;
            if(mixed) return launch_spmv_f64x(L, PREC_MIXED, P->p, P->Ap, dot, P->s, nullptr);
            return launch_spmv_f64x(L, PREC_F64, (level == 0) ? P->p : (const double*)V[level].r,
                                    (level == 0) ? P->Ap : (double*)V[level].x, dot, P->s, nullptr);
        }
        case 1: return launch_symgs(L, pr, r, r_f64, x, V[level].t, P->s, nullptr);
        case 2: return launch_symgs_zero(L, pr, r, r_f64, x, P->s, nullptr);
        case 3: return mg(*P, level, pr, r, r_f64, x, nullptr);
        case 4: case 5: case 6: case 7:
        {
            // the variants the V-cycle runs on this level
            int part = op - 4;
            if(L.u_ok && level + 1 < P->nlevels && (part == 1 || part == 2)) part += 3;
            return launch_sweep(L, pr, part, r, r_f64, x, V[level].t, P->s, nullptr);
        }
        case 8: return launch_restrict(L, pr, r, r_f64, x, V[level + 1].r, P->s, nullptr);
        case 9: return launch_read_stream(L, pr, P->partial, P->partial_cap, P->s);
        }
        return -2;
    };
    HPCGX_TRY(run());   // warm-up (and first-launch costs)
    hipEvent_t e0, e1;
    HPCGX_HIP(hipEventCreate(&e0));
    HPCGX_HIP(hipEventCreate(&e1));
    HPCGX_HIP(hipEventRecord(e0, P->s));
    for(int i = 0; i < reps; ++i) HPCGX_TRY(run());
    HPCGX_HIP(hipEventRecord(e1, P->s));
    HPCGX_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    HPCGX_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *avg_ms = ms / reps;
    return 0;
}

int hpcgx_set_shape(hpcgx_problem* p, int prec, int level, int op, int block, int rows)
{
    CHECK_P(p);
    if(prec < 0 || prec > 1 || level < -1 || level >= p->P->nlevels || op < -1 || op >= SH_COUNT ||
       !shape_supported(rows, block))
    {
        set_error("invalid precision, level, operator or launch shape");
        return -2;
    }
    for(int l = 0; l < p->P->nlevels; ++l)
    {
        if(level >= 0 && l != level) continue;
        for(int o = 0; o < SH_COUNT; ++o)
        {
            if(op >= 0 && o != op) continue;
            p->P->L[l].shp[prec][o].block = block;
            p->P->L[l].shp[prec][o].rows = (o == SH_RESTRICT) ? 1 : rows;
        }
    }
    p->P->invalidate_graphs();
    return 0;
}

int hpcgx_set_sweep_columns(hpcgx_problem* p, int level, int columns)
{
    CHECK_P(p);
    if(columns < 0 || level < -1 || level >= p->P->nlevels)
    {
        set_error("invalid level or column count");
        return -2;
    }
    for(int l = 0; l < p->P->nlevels; ++l)
        if(level < 0 || l == level) p->P->L[l].zcols = columns;
    // a captured CG iteration keeps the block order it was captured with
    p->P->invalidate_graphs();
    return 0;
}

int hpcgx_get_shape(const hpcgx_problem* p, int prec, int level, int op, int* block, int* rows)
{
    CHECK_P(p);
    CHECK_LEVEL(p, level);
    CHECK_PTRS(block, rows);
    if(op < 0 || op >= SH_COUNT || prec < 0 || prec > 1)
    {
        set_error("invalid precision or operator");
        return -2;
    }
    *block = p->P->L[level].shp[prec][op].block;
    *rows = p->P->L[level].shp[prec][op].rows;
    return 0;
}

int hpcgx_synchronize(hpcgx_problem* p)
{
    CHECK_P(p);
    HPCGX_HIP(hipStreamSynchronize(p->P->s));
    return 0;
}

int hpcgx_malloc(void** ptr, size_t bytes)
{
    CHECK_PTRS(ptr);
    HPCGX_HIP(hipMalloc(ptr, bytes ? bytes : 8));
    return 0;
}

int hpcgx_free(void* ptr)
{
    if(ptr) HPCGX_HIP(hipFree(ptr));
    return 0;
}

int hpcgx_memcpy_h2d(void* dst, const void* src, size_t bytes)
{
    if(bytes == 0) return 0;
    CHECK_PTRS(dst, src);
    // solver streams are non-blocking: finish their queued work first
    HPCGX_HIP(hipDeviceSynchronize());
    HPCGX_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return 0;
}

int hpcgx_memcpy_d2h(void* dst, const void* src, size_t bytes)
{
    if(bytes == 0) return 0;
    CHECK_PTRS(dst, src);
    HPCGX_HIP(hipDeviceSynchronize());
    HPCGX_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return 0;
}

int hpcgx_memcpy_d2d(void* dst, const void* src, size_t bytes)
{
    if(bytes == 0) return 0;
    CHECK_PTRS(dst, src);
    HPCGX_HIP(hipDeviceSynchronize());
    HPCGX_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
    HPCGX_HIP(hipDeviceSynchronize());
    return 0;
}

int hpcgx_device_count(int* count)
{
    CHECK_PTRS(count);
    *count = 0;
    HPCGX_HIP(hipGetDeviceCount(count));
    return 0;
}

int hpcgx_batch_create(int nprob, const int* dims, int nlevels, int stores, void* stream, hpcgx_batch** out)
{
    if(!out || !dims)
    {
        set_error("null argument");
        return -2;
    }
    *out = nullptr;
    Batch* B = nullptr;
    HPCGX_TRY(batch_create(B, nprob, dims, nlevels, (stores & HPCGX_STORE_F64) != 0,
                           (stores & HPCGX_STORE_MIXED) != 0, (hipStream_t)stream));
    *out = new hpcgx_batch{B};
    return 0;
}

int hpcgx_batch_destroy(hpcgx_batch* b)
{
    if(!b) return 0;
    batch_destroy(b->B);
    delete b;
    return 0;
}

int hpcgx_batch_problem(hpcgx_batch* b, int k, hpcgx_problem** out)
{
    if(!b || !b->B || !out || k < 0 || k >= (int)b->B->probs.size())
    {
        set_error("invalid batch or problem index");
        return -2;
    }
    *out = new hpcgx_problem{b->B->probs[k], false};
    return 0;
}

int hpcgx_batch_set_block(hpcgx_batch* b, int level, int block)
{
    if(!b || !b->B)
    {
        set_error("null batch");
        return -2;
    }
    return batch_set_block(*b->B, level, block);
}

int hpcgx_batch_set_rows(hpcgx_batch* b, int level, int rows)
{
    if(!b || !b->B)
    {
        set_error("null batch");
        return -2;
    }
    return batch_set_rows(*b->B, level, rows);
}

int hpcgx_batch_cg(hpcgx_batch* b, int mode, int max_iter, double tolerance, int use_graph, int* niters,
                   double* normr, double* normr0)
{
    if(!b || !b->B)
    {
        set_error("null batch");
        return -2;
    }
    return batch_cg(*b->B, mode, max_iter, tolerance, use_graph != 0, niters, normr, normr0);
}

int hpcgx_batch_profile(hpcgx_batch* b, int mode, int iters, double* kind_ms, double* kind_bytes, double* iter_ms)
{
    if(!b || !b->B)
    {
        set_error("null batch");
        return -2;
    }
    CHECK_PTRS(kind_ms, kind_bytes, iter_ms);
    return batch_profile(*b->B, mode, iters, kind_ms, kind_bytes, iter_ms);
}

int hpcgx_batch_history(hpcgx_batch* b, int k, double* history, int hist_len, int* count)
{
    if(!b || !b->B)
    {
        set_error("null batch");
        return -2;
    }
    CHECK_PTRS(history, count);
    if(k < 0 || k >= (int)b->B->probs.size() || hist_len < 1)
    {
        set_error("problem index or history length out of range");
        return -2;
    }
    Problem* P = b->B->probs[k];
    HPCGX_HIP(hipStreamSynchronize(b->B->s));
    CGScalars h;
    HPCGX_HIP(hipMemcpy(&h, P->sc, sizeof(h), hipMemcpyDeviceToHost));
    int n = std::min(h.iter + 1, std::min(hist_len, P->hist_cap + 1));
    HPCGX_HIP(hipMemcpy(history, P->hist, sizeof(double) * n, hipMemcpyDeviceToHost));
    *count = n;
    return 0;
}

const char* hpcgx_last_error(void) { return last_error(); }

const char* hpcgx_version(void) { return "rochpcg_amd 0.1 (gfx950)"; }

}  // extern "C"
