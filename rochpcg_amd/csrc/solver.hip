// solver.hip -- hierarchy construction, the multigrid V-cycle and the CG
// driver with device-resident scalars.
//
// Reference behaviour:
//   V-cycle       ComputeMG.cpp:64-102 (zero-guess pre-smoother, fused
//                 residual+restriction, recursion, prolongation, one
//                 post-smoother; the coarsest level runs one zero-guess SymGS)
//   CG            CG.cpp:104-200 (fused r update and ||r||, x update)
//   hierarchy     GenerateCoarseProblem.cpp:98-184, OptimizeProblem.cpp:69-160
//
// The CG keeps alpha, beta, rtz and ||r|| on the device: no host round trip
// per iteration.  A `done` flag written by the ||r|| reduction gates every
// kernel of later iterations, so iterations can be enqueued (or replayed as a
// captured hipGraph) ahead of the convergence test.
#include "refpath.hpp"
#include "setup.hpp"
#include "solver.hpp"

#include <cmath>
#include <cstring>

namespace hpcgx
{

static int alloc(void** p, size_t bytes)
{
    if(bytes == 0) bytes = 8;
    HPCGX_HIP(hipMalloc(p, bytes));
    return 0;
}

int problem_create(Problem*& out, int nx, int ny, int nz, int nlevels, bool with_f64, bool with_mixed, hipStream_t s)
{
    out = nullptr;
    if(nlevels < 1 || nlevels > MAX_LEVELS || nx < 2 || ny < 2 || nz < 2)
    {
        set_error("invalid problem dimensions or level count");
        return -2;
    }
    if(!with_f64 && !with_mixed)
    {
        set_error("at least one precision must be requested");
        return -2;
    }
    for(int l = 0; l + 1 < nlevels; ++l)
    {
        // GenerateCoarseProblem.cpp:106-109: every fine dimension must be even
        if(((nx >> l) & 1) || ((ny >> l) & 1) || ((nz >> l) & 1) || (nx >> (l + 1)) < 1 || (ny >> (l + 1)) < 1 ||
           (nz >> (l + 1)) < 1)
        {
            set_error("grid dimensions must be divisible by 2^(levels-1)");
            return -2;
        }
    }
    Problem* P = new Problem;
    P->nlevels = nlevels;
    P->has_f64 = with_f64;
    P->has_mixed = with_mixed;
    if(s)
        P->s = s;
    else
    {
        hipError_t e = hipStreamCreateWithFlags(&P->s, hipStreamNonBlocking);
        if(e != hipSuccess)
        {
            set_error(std::string("hipStreamCreateWithFlags: ") + hipGetErrorString(e));
            delete P;
            return -1;
        }
        P->own_stream = true;
    }
    auto fail = [&](int rc) {
        problem_destroy(P);
        return rc;
    };
    for(int l = 0; l < nlevels; ++l)
    {
        int rc = build_level(P->L[l], nx >> l, ny >> l, nz >> l, with_f64, with_mixed, P->s);
        if(rc) return fail(rc);
    }
    for(int l = 0; l + 1 < nlevels; ++l)
    {
        int rc = build_transfer(P->L[l], P->L[l + 1], P->s);
        if(rc) return fail(rc);
    }
    for(int l = 0; l < nlevels; ++l)
    {
        const size_t m = (size_t)P->L[l].m;
        if(with_f64)
        {
            if(alloc(&P->v64[l].t, m * 8)) return fail(-1);
            if(l > 0 && (alloc(&P->v64[l].r, m * 8) || alloc(&P->v64[l].x, m * 8))) return fail(-1);
        }
        if(with_mixed)
        {
            if(alloc(&P->v32[l].t, m * 4)) return fail(-1);
            if(l > 0 && (alloc(&P->v32[l].r, m * 4) || alloc(&P->v32[l].x, m * 4))) return fail(-1);
        }
    }
    const size_t m0 = (size_t)P->L[0].m;
    double** vecs[] = {&P->b, &P->x, &P->xexact, &P->r, &P->p, &P->Ap, &P->tmp};
    for(double** v : vecs)
        if(alloc((void**)v, m0 * 8)) return fail(-1);
    if(with_f64 && alloc((void**)&P->z64, m0 * 8)) return fail(-1);
    if(with_mixed && alloc((void**)&P->z32, m0 * 4)) return fail(-1);
    P->hist_cap = 4096;
    // a batched fused SpMV + dot keeps one partial per 64-row block
    P->partial_cap = std::max<int64_t>(FOLD_MAX, (int64_t)(m0 / 64) + 64);
    if(alloc((void**)&P->sc, sizeof(CGScalars)) || alloc((void**)&P->hist, sizeof(double) * (P->hist_cap + 1)) ||
       alloc((void**)&P->partial, sizeof(double) * P->partial_cap) ||
       alloc((void**)&P->dout, 8 * sizeof(double)))
        return fail(-1);
    if(hipMemsetAsync(P->sc, 0, sizeof(CGScalars), P->s) != hipSuccess)
        return fail(-1);
    if(make_rhs(P->L[0], 0, 0, P->b, P->xexact, P->x, P->s)) return fail(-1);
    if(hipStreamSynchronize(P->s) != hipSuccess)
    {
        set_error("stream synchronisation failed after setup");
        return fail(-1);
    }
    out = P;
    return 0;
}

void problem_destroy(Problem* P)
{
    if(!P) return;
    if(P->s) (void)hipStreamSynchronize(P->s);
    refpath_free(*P);
    for(int g = 0; g < 2; ++g)
    {
        if(P->gexec[g]) (void)hipGraphExecDestroy(P->gexec[g]);
        if(P->graph[g]) (void)hipGraphDestroy(P->graph[g]);
    }
    auto f = [](void* p) {
        if(p) (void)hipFree(p);
    };
    for(int l = 0; l < MAX_LEVELS; ++l)
    {
        free_level(P->L[l]);
        f(P->v64[l].r);
        f(P->v64[l].x);
        f(P->v64[l].t);
        f(P->v32[l].r);
        f(P->v32[l].x);
        f(P->v32[l].t);
    }
    f(P->b);
    f(P->x);
    f(P->xexact);
    if(!P->ext_cg)
    {
        f(P->r);
        f(P->p);
        f(P->Ap);
        f(P->z64);
    }
    f(P->z32);
    f(P->tmp);
    f(P->sc);
    f(P->hist);
    f(P->partial);
    f(P->dout);
    if(P->own_stream && P->s) (void)hipStreamDestroy(P->s);
    delete P;
}

int adopt_cg_vectors(Problem& P, double* r, double* z, double* p, double* Ap)
{
    if(!r || !z || !p || !Ap)
    {
        set_error("null CG vector");
        return -2;
    }
    HPCGX_HIP(hipStreamSynchronize(P.s));
    if(!P.ext_cg)
    {
        (void)hipFree(P.r);
        (void)hipFree(P.z64);
        (void)hipFree(P.p);
        (void)hipFree(P.Ap);
    }
    P.r = r;
    P.z64 = z;
    P.p = p;
    P.Ap = Ap;
    P.ext_cg = true;
    P.invalidate_graphs();
    return 0;
}

// ---------------------------------------------------------------------------
int mg(Problem& P, int l, Prec prec, const void* r, bool r_f64, void* x, const int* gate)
{
    Level& F = P.L[l];
    LevelVecs* V = (prec == PREC_F64) ? P.v64 : P.v32;
    if(l + 1 < P.nlevels)
    {
        LevelVecs& C = V[l + 1];
        // V[l].t carries the pre-smoother's upper sums to the post-smoother
        // when the correction touches only last-colour rows (F.u_ok)
        HPCGX_TRY(launch_symgs_zero(F, prec, r, r_f64, x, P.s, gate, F.u_ok ? V[l].t : nullptr));
        HPCGX_TRY(launch_restrict(F, prec, r, r_f64, x, C.r, P.s, gate));
        HPCGX_TRY(mg(P, l + 1, prec, C.r, prec == PREC_F64, C.x, gate));
        HPCGX_TRY(launch_prolong(F, prec, C.x, x, P.s, gate));
        HPCGX_TRY(launch_symgs(F, prec, r, r_f64, x, V[l].t, P.s, gate, F.u_ok));
    }
    else
    {
        HPCGX_TRY(launch_symgs_zero(F, prec, r, r_f64, x, P.s, gate));
    }
    return 0;
}

// ---------------------------------------------------------------------------
static RedArgs red(Problem& P, int op)
{
    RedArgs ra;
    ra.partial = P.partial;
    ra.op = op;
    ra.sc = P.sc;
    ra.hist = P.hist;
    ra.hist_cap = P.hist_cap;
    ra.out = P.dout;
    ra.cap = P.partial_cap;
    return ra;
}

static Prec outer_prec(const Problem& P) { return (P.mode == CG_MIXED) ? PREC_MIXED : PREC_F64; }

int cg_begin(Problem& P, int mode, bool precond, const double* b, double* x, int max_iter, double tol)
{
    if(mode == CG_F64 && !P.has_f64)
    {
        set_error("problem was built without f64 values");
        return -2;
    }
    if(mode == CG_MIXED && !P.has_mixed)
    {
        set_error("problem was built without the mixed-precision store");
        return -2;
    }
    if(max_iter < 0 || max_iter > P.hist_cap)
    {
        set_error("max_iter out of range (history capacity " + std::to_string(P.hist_cap) + ")");
        return -2;
    }
    P.mode = mode;
    P.precond = precond;
    P.host_iter = 0;
    P.cur_b = b ? b : P.b;
    P.cur_x = x ? x : P.x;
    CGScalars& h = P.h_sc;
    std::memset(&h, 0, sizeof(h));
    h.tol = tol;
    h.max_iter = max_iter;
    HPCGX_HIP(hipMemcpyAsync(P.sc, &h, sizeof(h), hipMemcpyHostToDevice, P.s));
    HPCGX_HIP(hipStreamSynchronize(P.s));
    const int64_t m = P.L[0].m;
    // CG.cpp:133-137: p = x; Ap = A p; r = b - Ap; normr = ||r||
    HPCGX_TRY(launch_copy(m, P.cur_x, true, P.p, true, P.s, nullptr));
    HPCGX_TRY(launch_spmv_f64x(P.L[0], outer_prec(P), P.p, P.Ap, nullptr, P.s, nullptr));
    HPCGX_TRY(launch_waxpby(m, 1.0, P.cur_b, -1.0, P.Ap, P.r, P.s, nullptr));
    RedArgs ra = red(P, OP_NORMR0);
    HPCGX_TRY(launch_dot(m, P.r, true, P.r, true, ra, P.s, nullptr));
    return 0;
}

static int enqueue_iteration(Problem& P, bool replace)
{
    const int* gate = &P.sc->done;
    const int64_t m = P.L[0].m;
    const bool mixed = P.mode == CG_MIXED;
    void* z = mixed ? (void*)P.z32 : (void*)P.z64;
    // CG.cpp:150-171
    if(P.precond)
        HPCGX_TRY(mg(P, 0, mixed ? PREC_MIXED : PREC_F64, P.r, true, z, gate));
    else
        HPCGX_TRY(launch_copy(m, P.r, true, z, !mixed, P.s, gate));
    HPCGX_TRY(launch_dot(m, P.r, true, z, !mixed, red(P, OP_RTZ), P.s, gate));
    HPCGX_TRY(launch_cg_p(m, z, !mixed, P.p, P.cur_x, P.sc, P.s, gate));
    RedArgs rp = red(P, OP_PAP);
    HPCGX_TRY(launch_spmv_f64x(P.L[0], outer_prec(P), P.p, P.Ap, &rp, P.s, gate));
    HPCGX_TRY(launch_cg_xr(m, P.Ap, P.r, red(P, OP_NORMR), P.s, gate));
    if(replace)
    {
        // true residual r = b - A x in f64 (iterative refinement of the recursive residual)
        HPCGX_TRY(launch_cg_xflush(m, P.p, P.cur_x, P.sc, P.s, gate));
        HPCGX_TRY(launch_spmv_f64x(P.L[0], outer_prec(P), P.cur_x, P.tmp, nullptr, P.s, gate));
        HPCGX_TRY(launch_waxpby(m, 1.0, P.cur_b, -1.0, P.tmp, P.r, P.s, gate));
        HPCGX_TRY(launch_dot(m, P.r, true, P.r, true, red(P, OP_NORMR_SET), P.s, gate));
    }
    return 0;
}

static bool replace_at(const Problem& P, int iter_index)
{
    return P.mode == CG_MIXED && P.replace_every > 0 && ((iter_index + 1) % P.replace_every) == 0;
}

static int ensure_graph(Problem& P, int g)
{
    const bool stale = P.gexec[g] == nullptr || P.graph_mode != P.mode || P.graph_precond != P.precond ||
                       P.graph_b != P.cur_b || P.graph_x != P.cur_x || P.graph_replace != P.replace_every;
    if(!stale) return 0;
    if(P.graph_mode != P.mode || P.graph_precond != P.precond || P.graph_b != P.cur_b || P.graph_x != P.cur_x ||
       P.graph_replace != P.replace_every)
    {
        for(int k = 0; k < 2; ++k)
        {
            if(P.gexec[k]) (void)hipGraphExecDestroy(P.gexec[k]);
            if(P.graph[k]) (void)hipGraphDestroy(P.graph[k]);
            P.gexec[k] = nullptr;
            P.graph[k] = nullptr;
        }
        P.graph_mode = P.mode;
        P.graph_precond = P.precond;
        P.graph_b = P.cur_b;
        P.graph_x = P.cur_x;
        P.graph_replace = P.replace_every;
    }
    HPCGX_HIP(hipStreamBeginCapture(P.s, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_iteration(P, g == 1);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(P.s, &graph);
    if(rc || e != hipSuccess)
    {
        if(graph) (void)hipGraphDestroy(graph);
        if(rc) return rc;
        set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        return -1;
    }
    e = hipGraphInstantiate(&P.gexec[g], graph, nullptr, nullptr, 0);
    if(e != hipSuccess)
    {
        (void)hipGraphDestroy(graph);
        P.gexec[g] = nullptr;
        set_error(std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        return -1;
    }
    P.graph[g] = graph;
    return 0;
}

int cg_iterations(Problem& P, int count, bool use_graph)
{
    for(int i = 0; i < count; ++i)
    {
        const int g = replace_at(P, P.host_iter) ? 1 : 0;
        if(use_graph)
        {
            HPCGX_TRY(ensure_graph(P, g));
            HPCGX_HIP(hipGraphLaunch(P.gexec[g], P.s));
        }
        else
        {
            HPCGX_TRY(enqueue_iteration(P, g == 1));
        }
        ++P.host_iter;
    }
    return 0;
}

int cg_iterations_profiled(Problem& P, int count, double* phase_ms)
{
    // events bracket each phase of every iteration (no host synchronisation
    // inside the loop); the sums are read once at the end
    const int per = 2 * PH_COUNT;
    std::vector<hipEvent_t> ev((size_t)count * per, nullptr);
    for(auto& e : ev) HPCGX_HIP(hipEventCreate(&e));
    const int* gate = &P.sc->done;
    const int64_t m = P.L[0].m;
    const bool mixed = P.mode == CG_MIXED;
    void* z = mixed ? (void*)P.z32 : (void*)P.z64;
    std::vector<char> used((size_t)count * PH_COUNT, 0);
    int rc = 0;
    for(int i = 0; i < count && rc == 0; ++i)
    {
        hipEvent_t* E = ev.data() + (size_t)i * per;
        char* U = used.data() + (size_t)i * PH_COUNT;
        auto beg = [&](int ph) {
            U[ph] = 1;
            return hipEventRecord(E[2 * ph], P.s);
        };
        auto end = [&](int ph) { return hipEventRecord(E[2 * ph + 1], P.s); };
        (void)beg(PH_MG);
        if(P.precond)
            rc = mg(P, 0, mixed ? PREC_MIXED : PREC_F64, P.r, true, z, gate);
        else
            rc = launch_copy(m, P.r, true, z, !mixed, P.s, gate);
        (void)end(PH_MG);
        if(rc) break;
        (void)beg(PH_DOT);
        rc = launch_dot(m, P.r, true, z, !mixed, red(P, OP_RTZ), P.s, gate);
        (void)end(PH_DOT);
        if(rc) break;
        (void)beg(PH_WAXPBY);
        rc = launch_cg_p(m, z, !mixed, P.p, P.cur_x, P.sc, P.s, gate);
        if(rc) break;
        (void)beg(PH_SPMV);
        RedArgs rp = red(P, OP_PAP);
        rc = launch_spmv_f64x(P.L[0], outer_prec(P), P.p, P.Ap, &rp, P.s, gate);
        (void)end(PH_SPMV);
        if(rc) break;
        rc = launch_cg_xr(m, P.Ap, P.r, red(P, OP_NORMR), P.s, gate);
        (void)end(PH_WAXPBY);
        if(rc) break;
        if(replace_at(P, P.host_iter))
        {
            (void)beg(PH_REPLACE);
            rc = launch_cg_xflush(m, P.p, P.cur_x, P.sc, P.s, gate);
            if(!rc) rc = launch_spmv_f64x(P.L[0], outer_prec(P), P.cur_x, P.tmp, nullptr, P.s, gate);
            if(!rc) rc = launch_waxpby(m, 1.0, P.cur_b, -1.0, P.tmp, P.r, P.s, gate);
            if(!rc) rc = launch_dot(m, P.r, true, P.r, true, red(P, OP_NORMR_SET), P.s, gate);
            (void)end(PH_REPLACE);
        }
        ++P.host_iter;
    }
    if(rc == 0 && hipStreamSynchronize(P.s) != hipSuccess)
    {
        set_error("stream synchronisation failed in a profiled CG run");
        rc = -1;
    }
    for(int ph = 0; ph < PH_COUNT; ++ph) phase_ms[ph] = 0.0;
    if(rc == 0)
    {
        for(int i = 0; i < count; ++i)
            for(int ph = 0; ph < PH_COUNT; ++ph)
            {
                if(!used[(size_t)i * PH_COUNT + ph]) continue;
                float ms = 0.f;
                hipEvent_t* E = ev.data() + (size_t)i * per;
                if(hipEventElapsedTime(&ms, E[2 * ph], E[2 * ph + 1]) == hipSuccess) phase_ms[ph] += ms;
            }
        // the waxpby phase brackets the SpMV too; report it without
        phase_ms[PH_WAXPBY] -= phase_ms[PH_SPMV];
    }
    for(auto& e : ev)
        if(e) (void)hipEventDestroy(e);
    return rc;
}

int cg_end(Problem& P, int* niters, double* normr, double* normr0, double* history, int hist_len)
{
    // the last iteration's x += alpha*p (deferred to the next p update, which
    // did not come), once: clear the flag so that further iterations or a
    // second cg_end do not apply it again
    if(P.cur_x)
    {
        HPCGX_TRY(launch_cg_xflush(P.L[0].m, P.p, P.cur_x, P.sc, P.s, nullptr));
        HPCGX_HIP(hipMemsetAsync(&P.sc->xpend, 0, sizeof(int), P.s));
    }
    HPCGX_HIP(hipStreamSynchronize(P.s));
    CGScalars h;
    HPCGX_HIP(hipMemcpy(&h, P.sc, sizeof(h), hipMemcpyDeviceToHost));
    if(niters) *niters = h.iter;
    if(normr) *normr = h.normr;
    if(normr0) *normr0 = h.normr0;
    if(history && hist_len > 0)
    {
        int n = h.iter + 1;
        if(n > hist_len) n = hist_len;
        if(n > P.hist_cap + 1) n = P.hist_cap + 1;
        HPCGX_HIP(hipMemcpy(history, P.hist, sizeof(double) * n, hipMemcpyDeviceToHost));
    }
    return 0;
}

int cg_solve(Problem& P, int mode, bool precond, const double* b, double* x, int max_iter, double tol, bool use_graph,
             int* niters, double* normr, double* normr0, double* history, int hist_len)
{
    HPCGX_TRY(cg_begin(P, mode, precond, b, x, max_iter, tol));
    const int check = 10;   // host looks at the convergence flag every `check` iterations
    while(P.host_iter < max_iter)
    {
        int n = max_iter - P.host_iter;
        if(tol > 0.0 && n > check) n = check;
        HPCGX_TRY(cg_iterations(P, n, use_graph));
        if(tol > 0.0 && P.host_iter < max_iter)
        {
            int done = 0;
            HPCGX_HIP(hipMemcpyAsync(&done, &P.sc->done, sizeof(int), hipMemcpyDeviceToHost, P.s));
            HPCGX_HIP(hipStreamSynchronize(P.s));
            if(done) break;
        }
    }
    return cg_end(P, niters, normr, normr0, history, hist_len);
}

// Algorithmic HBM bytes of one V-cycle from level l down, as this library's
// kernels run it: every slot a kernel loads once (4 B column + value; a sweep
// loads its colour's window bucket for each row), neighbour x
// gathers served from cache, own-row vectors read and written once.
static double vcycle_bytes(const Problem& P, int l, bool mixed)
{
    const Level& L = P.L[l];
    const double ib = 4.0, vb = mixed ? 2.0 : 8.0, xb = mixed ? 4.0 : 8.0;
    const double rb = (l == 0) ? 8.0 : xb;   // the CG residual is f64
    const double slot = ib + vb;
    const int last = L.ncolors - 1;
    const bool coarsest = l + 1 >= P.nlevels;
    const bool keep_u = !coarsest && L.u_ok;
    double b = 0.0;
    for(int c = 0; c <= last; ++c)
    {
        const double n = L.sizes[c];
        b += window_slots(L.win[WIN_LOWER][c]) * n * slot + n * (4 + vb + rb + xb);   // fwd_zero
        if(c < last) b += window_slots(L.win[WIN_UPPER][c]) * n * slot + n * (4 + vb + 2 * xb + (keep_u ? xb : 0));   // bwd_zero
    }
    if(coarsest) return b;
    const double mc = (double)L.mc;
    b += mc * (W * slot + rb + 4 + 4 + xb);   // fused residual + restriction (whole rows)
    b += vcycle_bytes(P, l + 1, mixed);
    b += mc * (4 + 4 + 3 * xb);                // prolongation
    for(int c = 0; c <= last; ++c)
    {
        const double n = L.sizes[c];
        if(c < last && keep_u)
        {
            int na = 0, nb = 0;
            window_pair_slots(L.win[WIN_LOWER][c], L.win[WIN_TAIL][c], &na, &nb);
            b += (na + nb) * n * slot + n * (4 + vb + rb + 3 * xb);
        }
        else
            b += n * (W * slot + rb + xb + (c < last ? xb : 0));   // whole-row forward sweep
        if(c < last) b += window_slots(L.win[WIN_UPPER][c]) * n * slot + n * (4 + vb + 2 * xb);   // bwd_full
    }
    return b;
}

int iteration_bytes(const Problem& P, int mode, double* vcycle, double* iteration)
{
    const bool mixed = mode == CG_MIXED;
    const double mg = vcycle_bytes(P, 0, mixed);
    const double m0 = (double)P.L[0].m;
    const double vb = mixed ? 2.0 : 8.0, zb = mixed ? 4.0 : 8.0;
    double it = mg;
    it += m0 * (W * (4 + vb) + 16);            // SpMV + p'Ap
    it += m0 * (8 + zb);                       // r'z
    it += m0 * (zb + 32);                      // x += alpha p (deferred), p = z + beta p
    it += m0 * 24;                             // r update and ||r||
    if(mixed && P.replace_every > 0)           // true residual every k iterations
        it += m0 * (W * (4 + vb) + 16 + 24 + 16) / P.replace_every;
    if(vcycle) *vcycle = mg;
    if(iteration) *iteration = it;
    return 0;
}

int64_t problem_device_bytes(const Problem& P)
{
    int64_t b = 0;
    for(int l = 0; l < P.nlevels; ++l)
    {
        const Level& L = P.L[l];
        const int64_t m = L.m, me = ell_rows(m);
        b += 4 * W * me + 4 * 4 * m;   // columns; diag_idx, desc, perm, colors
        if(P.has_f64) b += 8 * W * me + 8 * m + 8 * m + (l > 0 ? 16 * m : 0);   // values, 1/a_ii, t, r, x
        if(P.has_mixed) b += 2 * W * me + 4 * m + 4 * m + (l > 0 ? 8 * m : 0);
        b += 3 * 4 * L.mc;   // transfer maps
    }
    const int64_t m0 = P.L[0].m;
    b += 7 * 8 * m0 + (P.has_f64 ? 8 * m0 : 0) + (P.has_mixed ? 4 * m0 : 0);   // CG vectors, z
    b += 8 * P.partial_cap + 8 * (P.hist_cap + 1) + (int64_t)sizeof(CGScalars) + 64;
    return b;
}

double hpcg_flops_per_iter(const Problem& P)
{
    // ReportResults.cpp:73-89, one iteration, 1 pre- and 1 post-smoother
    const double nrow = (double)P.L[0].m;
    double f = 12.0 * nrow + 2.0 * (double)P.L[0].nnz;
    for(int l = 0; l + 1 < P.nlevels; ++l) f += 10.0 * (double)P.L[l].nnz;
    f += 4.0 * (double)P.L[P.nlevels - 1].nnz;
    return f;
}

}  // namespace hpcgx
