// solver.hpp -- the multigrid hierarchy, the V-cycle and the device-scalar CG.
#pragma once
#include "common.hpp"
#include "ops.hpp"

namespace hpcgx
{

enum CGMode : int
{
    CG_F64 = 0,    // f64 everywhere (the reference's arithmetic)
    CG_MIXED = 1   // f64 outer CG; MG with f32 vectors and f16 ELL values
};

// Per-phase event timing of the CG iteration (profiled runs only).
enum Phase : int
{
    PH_SPMV = 0,       // outer SpMV (+ fused p'Ap)
    PH_MG = 1,         // whole V-cycle
    PH_DOT = 2,        // r'z
    PH_WAXPBY = 3,     // p update + fused x/r update and ||r||
    PH_REPLACE = 4,    // true-residual recomputation (mixed mode)
    PH_COUNT = 5
};

struct LevelVecs
{
    void* r = nullptr;   // level rhs (levels >= 1)
    void* x = nullptr;   // level solution (levels >= 1)
    void* t = nullptr;   // SymGS partial sums
};

struct RefPath;

struct Problem
{
    int nlevels = 0;
    Level L[MAX_LEVELS];
    bool has_f64 = false, has_mixed = false;
    hipStream_t s = nullptr;
    bool own_stream = false;
    LevelVecs v64[MAX_LEVELS], v32[MAX_LEVELS];

    // CG data in multicolour order
    double *b = nullptr, *x = nullptr, *xexact = nullptr;
    double *r = nullptr, *p = nullptr, *Ap = nullptr, *z64 = nullptr, *tmp = nullptr;
    float* z32 = nullptr;

    CGScalars* sc = nullptr;
    double* hist = nullptr;
    int hist_cap = 0;
    double* partial = nullptr;
    int64_t partial_cap = 0;   // reduction partial slots
    double* dout = nullptr;   // OP_STORE results
    bool ext_cg = false;      // r/z64/p/Ap belong to the caller (drop-in CGData)
    double h_out[8] = {};     // host landing slot of OP_STORE results
    CGScalars h_sc = {};      // host image of the scalars uploaded by cg_begin

    // CG run state
    int mode = CG_F64;
    bool precond = true;
    int replace_every = 0;    // mixed mode: true residual every k iterations (0 = never)
    int host_iter = 0;        // iterations enqueued since cg_begin
    const double* cur_b = nullptr;
    double* cur_x = nullptr;
    const double* graph_b = nullptr;
    double* graph_x = nullptr;
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    int graph_mode = -1;
    bool graph_precond = true;
    int graph_replace = -1;
    RefPath* ref = nullptr;   // natural-order work vectors of the reference path (refpath.hip)

    // forget captured graphs (launch shapes or operands changed)
    void invalidate_graphs()
    {
        for(int k = 0; k < 2; ++k)
        {
            if(gexec[k]) (void)hipGraphExecDestroy(gexec[k]);
            if(graph[k]) (void)hipGraphDestroy(graph[k]);
            gexec[k] = nullptr;
            graph[k] = nullptr;
        }
        graph_mode = -1;
    }
};

int problem_create(Problem*& P, int nx, int ny, int nz, int nlevels, bool with_f64, bool with_mixed, hipStream_t s);
void problem_destroy(Problem* P);
// use caller-owned CG vectors (CGData of the drop-in layer) instead of the problem's own
int adopt_cg_vectors(Problem& P, double* r, double* z, double* p, double* Ap);

// V-cycle on `level` (r is f64 at level 0 of both modes; x is f64 or f32)
int mg(Problem& P, int level, Prec prec, const void* r, bool r_f64, void* x, const int* gate);

// CG: preamble (x0 -> r, normr0), k iterations, results
int cg_begin(Problem& P, int mode, bool precond, const double* b, double* x, int max_iter, double tol);
int cg_iterations(Problem& P, int count, bool use_graph);
int cg_iterations_profiled(Problem& P, int count, double* phase_ms);
int cg_end(Problem& P, int* niters, double* normr, double* normr0, double* history, int hist_len);
// device bytes the problem holds (matrices, maps, vectors, scratch)
int64_t problem_device_bytes(const Problem& P);
// algorithmic HBM bytes of one V-cycle and of one CG iteration in `mode`
int iteration_bytes(const Problem& P, int mode, double* vcycle, double* iteration);
int cg_solve(Problem& P, int mode, bool precond, const double* b, double* x, int max_iter, double tol,
             bool use_graph, int* niters, double* normr, double* normr0, double* history, int hist_len);

double hpcg_flops_per_iter(const Problem& P);

}  // namespace hpcgx
