// batch.hip -- batched solves of many independent, unevenly sized HPCG
// problems.  Each problem keeps its own hierarchy, vectors, reduction scratch
// and CG scalars (a Problem); the V-cycle and the CG iteration are issued as
// one launch per (operator, level, colour) covering every problem, so a whole
// batched iteration is ~115 launches whatever the number of problems, and is
// captured once as a hipGraph.  Blocks map to (problem, row block) through a
// segment table in device memory (the segment records, then the segment of
// every block); a problem's segments return at once after its own
// convergence flag is set.
//
// Per-row arithmetic is shared with the single-problem kernels
// (kernels.hpp), so each problem computes exactly what it would alone, up to
// the grouping of the dot-product partials.
#include "batch.hpp"
#include "kernels.hpp"
#include "setup.hpp"

#include <algorithm>
#include <cstring>
#include <functional>

namespace hpcgx
{

// the fold of a reducing launch: block k folds segment k's partials (one per
// block of the segment) for that segment's problem
template <int BS>
__global__ __launch_bounds__(BS) void kb_fold(const BSeg* __restrict__ segs)
{
    const BSeg& S = segs[blockIdx.x];
    if(gated(S.gate) || S.ra.partial == nullptr) return;
    fold_partials<BS>(S.ra.partial, S.nblk, S.ra);
}

#define BSEG()                                                \
    const int64_t b_ = xcd_block(blockIdx.x, gridDim.x);      \
    const BSeg& S = segs[seg_map(segs, nseg)[b_]];            \
    if(gated(S.gate)) return;                                 \
    const int64_t lb = b_ - S.blk0;

template <typename VT, typename XT, typename YT, bool DOT, int BS>
__global__ __launch_bounds__(BS) void kb_spmv(const BSeg* __restrict__ segs, int nseg)
{
    BSEG();
    const int64_t i = lb * BS + threadIdx.x;
    double acc = 0.0;
    if(i < S.size)
    {
        const int64_t row = S.off + i;
        XT s[1];
        row_spmv<VT, XT, 1>(row, S.col, (const VT*)S.val, (const XT*)S.r, s);
        ((YT*)S.x)[row] = (YT)s[0];
        if(DOT) acc = fmaT((double)((const XT*)S.r)[row], (double)s[0], acc);
    }
    if(DOT) block_partial<BS>(acc, S.ra.partial, lb);
}

// The sweeps are compiled for the widest window of the launch's segments
// (BLaunch::win), like the single-problem kernels, and for R rows per thread
// (BLaunch::rows; 2 only when every segment's offset and size are even).
template <int NS, typename VT, typename XT, typename RT, int R, int BS>
__global__ __launch_bounds__(BS) void kb_fwd_zero(const BSeg* __restrict__ segs, int nseg)
{
    BSEG();
    const int64_t i = (lb * BS + threadIdx.x) * R;
    if(i >= S.size) return;
    row_fwd_zero<NS, VT, XT, RT, R>(S.off + i, S.col, (const VT*)S.val, (const VT*)S.dg, S.desc,
                                    (const RT*)S.r, (XT*)S.x);
}

template <int NS, typename VT, typename XT, int R, int BS>
__global__ __launch_bounds__(BS) void kb_bwd_zero(const BSeg* __restrict__ segs, int nseg)
{
    BSEG();
    const int64_t i = (lb * BS + threadIdx.x) * R;
    if(i >= S.size) return;
    if(S.t)
        row_bwd_zero<NS, VT, XT, true, R>(S.off + i, S.col, (const VT*)S.val, (const VT*)S.dg, S.desc,
                                          (XT*)S.x, (XT*)S.t);
    else
        row_bwd_zero<NS, VT, XT, false, R>(S.off + i, S.col, (const VT*)S.val, (const VT*)S.dg, S.desc,
                                           (XT*)S.x, nullptr);
}

// every segment reuses its pre-smoother's upper sums (BLaunch::all_u)
template <int NA, int NB, typename VT, typename XT, typename RT, int R, int BS>
__global__ __launch_bounds__(BS) void kb_fwd_full_u(const BSeg* __restrict__ segs, int nseg)
{
    BSEG();
    const int64_t i = (lb * BS + threadIdx.x) * R;
    if(i >= S.size) return;
    row_fwd_full_u<NA, NB, VT, XT, RT, R>(S.off + i, S.col, (const VT*)S.val, (const VT*)S.dg, S.desc,
                                          (const RT*)S.r, (XT*)S.x, (XT*)S.t);
}

template <typename VT, typename XT, typename RT, int R, int BS>
__global__ __launch_bounds__(BS) void kb_fwd_full(const BSeg* __restrict__ segs, int nseg)
{
    BSEG();
    const int64_t i = (lb * BS + threadIdx.x) * R;
    if(i >= S.size) return;
    if(S.use_u)
        row_fwd_full_u<W - 1, W - 1, VT, XT, RT, R>(S.off + i, S.col, (const VT*)S.val, (const VT*)S.dg,
                                                    S.desc, (const RT*)S.r, (XT*)S.x, (XT*)S.t);
    else if(S.t)
        row_fwd_full<VT, XT, RT, true, R>(S.off + i, S.col, (const VT*)S.val, (const RT*)S.r, (XT*)S.x,
                                          (XT*)S.t);
    else
        row_fwd_full<VT, XT, RT, false, R>(S.off + i, S.col, (const VT*)S.val, (const RT*)S.r, (XT*)S.x,
                                           nullptr);
}

template <int NS, typename VT, typename XT, int R, int BS>
__global__ __launch_bounds__(BS) void kb_bwd_full(const BSeg* __restrict__ segs, int nseg)
{
    BSEG();
    const int64_t i = (lb * BS + threadIdx.x) * R;
    if(i >= S.size) return;
    row_bwd_full<NS, VT, XT, R>(S.off + i, S.col, (const VT*)S.val, (const VT*)S.dg, S.desc,
                                (const XT*)S.t, (XT*)S.x);
}

template <typename VT, typename XT, typename RT, int BS>
__global__ __launch_bounds__(BS) void kb_restrict(const BSeg* __restrict__ segs, int nseg)
{
    BSEG();
    const int64_t k = lb * BS + threadIdx.x;
    if(k >= S.size) return;
    row_restrict<VT, XT, RT, XT>(k, S.col, (const VT*)S.val, S.f2c_fine, S.f2c_coarse, (const RT*)S.r,
                                 (const XT*)S.x, (XT*)S.rc);
}

template <typename XT, int BS>
__global__ __launch_bounds__(BS) void kb_prolong(const BSeg* __restrict__ segs, int nseg)
{
    BSEG();
    const int64_t k = lb * BS + threadIdx.x;
    if(k >= S.size) return;
    const int64_t f = S.f2c_fine[k];
    XT* x = (XT*)S.x;
    x[f] = x[f] + ((const XT*)S.rc)[S.f2c_coarse[k]];
}

// vector segments: block lb of the segment handles elements [lb*chunk, ...)
#define BCHUNK()                                                   \
    const int64_t chunk_ = (S.size + S.nblk - 1) / S.nblk;         \
    const int64_t e0 = lb * chunk_;                                \
    const int64_t e1 = (e0 + chunk_ < S.size) ? e0 + chunk_ : S.size;

template <typename XT, typename YT, int BS>
__global__ __launch_bounds__(BS) void kb_dot(const BSeg* __restrict__ segs, int nseg)
{
    BSEG();
    BCHUNK();
    const XT* a = (const XT*)S.a;
    const YT* bb = (const YT*)S.b;
    double acc = 0.0;
    for(int64_t i = e0 + threadIdx.x; i < e1; i += BS) acc = fmaT((double)bb[i], (double)a[i], acc);
    block_partial<BS>(acc, S.ra.partial, lb);
}

template <typename ZT, int BS>
__global__ __launch_bounds__(BS) void kb_cg_p(const BSeg* __restrict__ segs, int nseg)
{
    BSEG();
    BCHUNK();
    const ZT* z = (const ZT*)S.a;
    double* p = (double*)S.w;
    const bool first = S.ra.sc->iter == 0;
    const double beta = S.ra.sc->beta;
    for(int64_t i = e0 + threadIdx.x; i < e1; i += BS) p[i] = first ? (double)z[i] : fmaT(beta, p[i], (double)z[i]);
}

template <int BS>
__global__ __launch_bounds__(BS) void kb_cg_xr(const BSeg* __restrict__ segs, int nseg)
{
    BSEG();
    BCHUNK();
    const double* p = (const double*)S.a;
    const double* Ap = (const double*)S.b;
    double* x = (double*)S.x;
    double* r = (double*)S.w;
    const double alpha = S.ra.sc->alpha;
    double acc = 0.0;
    for(int64_t i = e0 + threadIdx.x; i < e1; i += BS)
    {
        const double rv = fmaT(-alpha, Ap[i], r[i]);
        r[i] = rv;
        x[i] = fmaT(alpha, p[i], x[i]);
        acc = fmaT(rv, rv, acc);
    }
    block_partial<BS>(acc, S.ra.partial, lb);
}

template <int BS>
__global__ __launch_bounds__(BS) void kb_copy(const BSeg* __restrict__ segs, int nseg)
{
    BSEG();
    BCHUNK();
    const double* a = (const double*)S.a;
    double* w = (double*)S.w;
    for(int64_t i = e0 + threadIdx.x; i < e1; i += BS) w[i] = a[i];
}

// w = a + beta*b (ComputeWAXPBY's alpha == 1 form)
template <int BS>
__global__ __launch_bounds__(BS) void kb_waxpby(const BSeg* __restrict__ segs, int nseg, double beta)
{
    BSEG();
    BCHUNK();
    const double* a = (const double*)S.a;
    const double* bb = (const double*)S.b;
    double* w = (double*)S.w;
    for(int64_t i = e0 + threadIdx.x; i < e1; i += BS) w[i] = fmaT(beta, bb[i], a[i]);
}
#undef BSEG
#undef BCHUNK

// ---------------------------------------------------------------------------
// launching a planned record

template <typename F>
static void with_rows(int rows, F&& f)
{
    if(rows == 2)
        f(std::integral_constant<int, 2>{});
    else
        f(std::integral_constant<int, 1>{});
}

template <int BS>
static void launch_bs(const BLaunch& L, hipStream_t s)
{
    const dim3 g((unsigned)L.nblocks), b(BS);
    const BSeg* T = L.segs;
    const int n = L.nseg;
    const bool f64 = L.prec == PREC_F64;
    switch(L.kind)
    {
    case BK_SPMV:
        // outer SpMV: f64 vectors over either store; MG-internal SpMV is not batched
        if(f64)
        {
            if(L.dot)
                hipLaunchKernelGGL((kb_spmv<double, double, double, true, BS>), g, b, 0, s, T, n);
            else
                hipLaunchKernelGGL((kb_spmv<double, double, double, false, BS>), g, b, 0, s, T, n);
        }
        else
        {
            if(L.dot)
                hipLaunchKernelGGL((kb_spmv<__half, double, double, true, BS>), g, b, 0, s, T, n);
            else
                hipLaunchKernelGGL((kb_spmv<__half, double, double, false, BS>), g, b, 0, s, T, n);
        }
        break;
    case BK_FWD_ZERO:
        with_rows(L.rows, [&](auto RR) {
            constexpr int R = decltype(RR)::value;
            with_window(L.win, [&](auto NN) {
                constexpr int NS = decltype(NN)::value;
                if(f64)
                    hipLaunchKernelGGL((kb_fwd_zero<NS, double, double, double, R, BS>), g, b, 0, s, T, n);
                else if(L.r_f64)
                    hipLaunchKernelGGL((kb_fwd_zero<NS, __half, float, double, R, BS>), g, b, 0, s, T, n);
                else
                    hipLaunchKernelGGL((kb_fwd_zero<NS, __half, float, float, R, BS>), g, b, 0, s, T, n);
            });
        });
        break;
    case BK_BWD_ZERO:
        with_rows(L.rows, [&](auto RR) {
            constexpr int R = decltype(RR)::value;
            with_window(L.win, [&](auto NN) {
                constexpr int NS = decltype(NN)::value;
                if(f64)
                    hipLaunchKernelGGL((kb_bwd_zero<NS, double, double, R, BS>), g, b, 0, s, T, n);
                else
                    hipLaunchKernelGGL((kb_bwd_zero<NS, __half, float, R, BS>), g, b, 0, s, T, n);
            });
        });
        break;
    case BK_FWD_FULL:
        with_rows(L.rows, [&](auto RR) {
            constexpr int R = decltype(RR)::value;
            if(L.all_u)
                with_window_pair(L.win, L.win_tail, [&](auto NA_, auto NB_) {
                    constexpr int NA = decltype(NA_)::value, NB = decltype(NB_)::value;
                    if(f64)
                        hipLaunchKernelGGL((kb_fwd_full_u<NA, NB, double, double, double, R, BS>), g, b, 0, s, T,
                                           n);
                    else if(L.r_f64)
                        hipLaunchKernelGGL((kb_fwd_full_u<NA, NB, __half, float, double, R, BS>), g, b, 0, s, T,
                                           n);
                    else
                        hipLaunchKernelGGL((kb_fwd_full_u<NA, NB, __half, float, float, R, BS>), g, b, 0, s, T, n);
                });
            else if(f64)
                hipLaunchKernelGGL((kb_fwd_full<double, double, double, R, BS>), g, b, 0, s, T, n);
            else if(L.r_f64)
                hipLaunchKernelGGL((kb_fwd_full<__half, float, double, R, BS>), g, b, 0, s, T, n);
            else
                hipLaunchKernelGGL((kb_fwd_full<__half, float, float, R, BS>), g, b, 0, s, T, n);
        });
        break;
    case BK_BWD_FULL:
        with_rows(L.rows, [&](auto RR) {
            constexpr int R = decltype(RR)::value;
            with_window(L.win, [&](auto NN) {
                constexpr int NS = decltype(NN)::value;
                if(f64)
                    hipLaunchKernelGGL((kb_bwd_full<NS, double, double, R, BS>), g, b, 0, s, T, n);
                else
                    hipLaunchKernelGGL((kb_bwd_full<NS, __half, float, R, BS>), g, b, 0, s, T, n);
            });
        });
        break;
    case BK_RESTRICT:
        if(f64)
            hipLaunchKernelGGL((kb_restrict<double, double, double, BS>), g, b, 0, s, T, n);
        else if(L.r_f64)
            hipLaunchKernelGGL((kb_restrict<__half, float, double, BS>), g, b, 0, s, T, n);
        else
            hipLaunchKernelGGL((kb_restrict<__half, float, float, BS>), g, b, 0, s, T, n);
        break;
    case BK_PROLONG:
        if(f64)
            hipLaunchKernelGGL((kb_prolong<double, BS>), g, b, 0, s, T, n);
        else
            hipLaunchKernelGGL((kb_prolong<float, BS>), g, b, 0, s, T, n);
        break;
    case BK_DOT:
        if(!L.bf32)
            hipLaunchKernelGGL((kb_dot<double, double, BS>), g, b, 0, s, T, n);
        else
            hipLaunchKernelGGL((kb_dot<double, float, BS>), g, b, 0, s, T, n);
        break;
    case BK_CG_P:
        if(!L.bf32)
            hipLaunchKernelGGL((kb_cg_p<double, BS>), g, b, 0, s, T, n);
        else
            hipLaunchKernelGGL((kb_cg_p<float, BS>), g, b, 0, s, T, n);
        break;
    case BK_CG_XR: hipLaunchKernelGGL((kb_cg_xr<BS>), g, b, 0, s, T, n); break;
    case BK_COPY: hipLaunchKernelGGL((kb_copy<BS>), g, b, 0, s, T, n); break;
    case BK_WAXPBY: hipLaunchKernelGGL((kb_waxpby<BS>), g, b, 0, s, T, n, L.beta); break;
    }
}

static int launch(const BLaunch& L, hipStream_t s)
{
    if(L.nblocks == 0) return 0;
    switch(L.block)
    {
    case 64: launch_bs<64>(L, s); break;
    case 128: launch_bs<128>(L, s); break;
    default: launch_bs<256>(L, s); break;
    }
    HPCGX_HIP(hipGetLastError());
    if(L.kind == BK_DOT || L.kind == BK_CG_XR || (L.kind == BK_SPMV && L.dot))
    {
        hipLaunchKernelGGL((kb_fold<FOLD_THREADS>), dim3((unsigned)L.nseg), dim3(FOLD_THREADS), 0, s, L.segs);
        HPCGX_HIP(hipGetLastError());
    }
    return 0;
}

// ---------------------------------------------------------------------------
// planning

struct Planner
{
    Batch& B;
    std::vector<BLaunch>* out;
    int prec;
    int next_bf32 = 0;   // the next vector launch reads an f32 z
    // slot windows of the next sweep launch (full rows unless set)
    int next_win = W, next_tail = W, next_all_u = 0;
    // rows per thread of the next sweep launch (2 needs even segment offsets and sizes)
    int next_rows = 1;

    // widest window of kind `wk` of colour c on level l over the problems
    // that have the colour
    int widest(int l, int wk, int c)
    {
        int w = 0;
        for(Problem* P : B.probs)
            if(c < P->L[l].ncolors) w = std::max<int>(w, P->L[l].win[wk][c]);
        return w;
    }

    // append a launch whose segments are built by `fill(k, seg) -> bool`
    template <typename Fill>
    int add(int kind, int block, bool vector_kind, int r_f64, int dot, double beta, Fill&& fill)
    {
        std::vector<BSeg> segs;
        std::vector<size_t> prob;
        for(size_t k = 0; k < B.probs.size(); ++k)
        {
            BSeg sg;
            std::memset(&sg, 0, sizeof(sg));
            sg.gate = &B.probs[k]->sc->done;
            if(!fill((int)k, sg) || sg.size <= 0) continue;
            segs.push_back(sg);
            prob.push_back(k);
        }
        const int rows = launch_rows(segs.data(), segs.size(), next_rows);
        const int64_t nb = assign_blocks(segs.data(), segs.size(), block, vector_kind, rows);
        for(size_t j = 0; j < segs.size(); ++j)
            if(segs[j].ra.partial && segs[j].nblk > B.probs[prob[j]]->partial_cap)
            {
                set_error("batched reduction exceeds the partial capacity");
                return -2;
            }
        BLaunch L;
        std::memset(&L, 0, sizeof(L));
        for(const BSeg& sg : segs)
        {
            L.nrows += sg.size;
            if(sg.t) L.any_t = 1;
        }
        L.kind = kind;
        L.prec = prec;
        L.r_f64 = r_f64;
        L.dot = dot;
        L.block = block;
        L.nseg = (int)segs.size();
        L.nblocks = nb;
        L.beta = beta;
        L.bf32 = next_bf32;
        L.win = next_win;
        L.win_tail = next_tail;
        L.all_u = next_all_u;
        L.rows = rows;
        next_bf32 = 0;
        next_win = next_tail = W;
        next_all_u = 0;
        next_rows = 1;
        if(!segs.empty())
        {
            if(segs.size() > MAX_SEGS)
            {
                set_error("too many problems in one batched launch");
                return -2;
            }
            std::vector<char> h(table_bytes(segs.size(), nb));
            std::memcpy(h.data(), segs.data(), sizeof(BSeg) * segs.size());
            block_map(segs.data(), segs.size(), reinterpret_cast<uint16_t*>(h.data() + sizeof(BSeg) * segs.size()));
            BSeg* d = nullptr;
            HPCGX_HIP(hipMalloc(&d, h.size()));
            B.tables.push_back(d);
            HPCGX_HIP(hipMemcpy(d, h.data(), h.size(), hipMemcpyHostToDevice));
            L.segs = d;
            out->push_back(L);
        }
        return 0;
    }

    RedArgs red(Problem* P, int op)
    {
        RedArgs ra;
        ra.partial = P->partial;
        ra.op = op;
        ra.sc = P->sc;
        ra.hist = P->hist;
        ra.hist_cap = P->hist_cap;
        ra.out = P->dout;
        ra.cap = P->partial_cap;
        return ra;
    }

    const void* vals(const Level& L) { return prec == PREC_F64 ? (const void*)L.ell_val : (const void*)L.ell_h; }
    const void* diags(const Level& L) { return prec == PREC_F64 ? (const void*)L.diag64 : (const void*)L.diag_h; }
    LevelVecs* vecs(Problem* P) { return prec == PREC_F64 ? P->v64 : P->v32; }
    void* z(Problem* P) { return prec == PREC_F64 ? (void*)P->z64 : (void*)P->z32; }

    int max_colors(int l)
    {
        int n = 0;
        for(Problem* P : B.probs) n = std::max(n, P->L[l].ncolors);
        return n;
    }

    // the V-cycle from level l (ComputeMG.cpp:64-102), all problems in lockstep
    int mg(int l)
    {
        const int bs = B.block[l];
        const bool last = (l + 1 == B.nlevels);
        const int r64 = (l == 0) ? 1 : (prec == PREC_F64);
        auto level_seg = [&](int k, BSeg& s, int c) {
            Problem* P = B.probs[k];
            const Level& L = P->L[l];
            if(c >= L.ncolors) return false;
            s.off = L.offsets[c];
            s.size = L.sizes[c];
            s.col = L.ell_col;
            s.val = vals(L);
            s.desc = L.desc;
            s.dg = diags(L);
            s.r = (l == 0) ? (const void*)P->r : vecs(P)[l].r;
            s.x = (l == 0) ? z(P) : vecs(P)[l].x;
            return true;
        };
        const int C = max_colors(l);
        for(int c = 0; c < C; ++c)
        {
            next_win = widest(l, WIN_LOWER, c);
            next_rows = B.rows[l];
            HPCGX_TRY(add(BK_FWD_ZERO, bs, false, r64, 0, 0.0, [&](int k, BSeg& s) { return level_seg(k, s, c); }));
        }
        // the pre-smoother keeps its upper sums for the post-smoother where
        // the correction changes only last-colour rows (Level::u_ok)
        for(int c = C - 2; c >= 0; --c)
        {
            next_win = widest(l, WIN_UPPER, c);
            next_rows = B.rows[l];
            HPCGX_TRY(add(BK_BWD_ZERO, bs, false, r64, 0, 0.0, [&](int k, BSeg& s) {
                if(!(c <= B.probs[k]->L[l].ncolors - 2 && level_seg(k, s, c))) return false;
                if(!last && B.probs[k]->L[l].u_ok) s.t = vecs(B.probs[k])[l].t;
                return true;
            }));
        }
        if(last) return 0;
        HPCGX_TRY(add(BK_RESTRICT, bs, false, r64, 0, 0.0, [&](int k, BSeg& s) {
            Problem* P = B.probs[k];
            const Level& L = P->L[l];
            s.size = L.mc;
            s.col = L.ell_col;
            s.val = vals(L);
            s.f2c_fine = L.f2c_fine;
            s.f2c_coarse = L.f2c_coarse;
            s.r = (l == 0) ? (const void*)P->r : vecs(P)[l].r;
            s.x = (l == 0) ? z(P) : vecs(P)[l].x;
            s.rc = vecs(P)[l + 1].r;
            return true;
        }));
        HPCGX_TRY(mg(l + 1));
        HPCGX_TRY(add(BK_PROLONG, bs, false, r64, 0, 0.0, [&](int k, BSeg& s) {
            Problem* P = B.probs[k];
            const Level& L = P->L[l];
            s.size = L.mc;
            s.f2c_fine = L.f2c_fine;
            s.f2c_coarse = L.f2c_coarse;
            s.rc = vecs(P)[l + 1].x;
            s.x = (l == 0) ? z(P) : vecs(P)[l].x;
            return true;
        }));
        for(int c = 0; c < C; ++c)
        {
            bool all_u = true;
            for(Problem* P : B.probs)
                if(c < P->L[l].ncolors) all_u = all_u && c < P->L[l].ncolors - 1 && P->L[l].u_ok;
            if(all_u)
            {
                next_all_u = 1;
                next_win = widest(l, WIN_LOWER, c);
                next_tail = widest(l, WIN_TAIL, c);
            }
            next_rows = B.rows[l];
            HPCGX_TRY(add(BK_FWD_FULL, bs, false, r64, 0, 0.0, [&](int k, BSeg& s) {
                if(!level_seg(k, s, c)) return false;
                const bool before_last = c < B.probs[k]->L[l].ncolors - 1;
                s.t = before_last ? vecs(B.probs[k])[l].t : nullptr;
                s.use_u = before_last && B.probs[k]->L[l].u_ok;
                return true;
            }));
        }
        for(int c = C - 2; c >= 0; --c)
        {
            next_win = widest(l, WIN_UPPER, c);
            next_rows = B.rows[l];
            HPCGX_TRY(add(BK_BWD_FULL, bs, false, r64, 0, 0.0, [&](int k, BSeg& s) {
                if(!(c <= B.probs[k]->L[l].ncolors - 2 && level_seg(k, s, c))) return false;
                s.t = vecs(B.probs[k])[l].t;
                return true;
            }));
        }
        return 0;
    }

    int outer_spmv(bool dot, bool from_x, bool into_tmp)
    {
        return add(BK_SPMV, B.block[0], false, 1, dot ? 1 : 0, 0.0, [&](int k, BSeg& s) {
            Problem* P = B.probs[k];
            const Level& L = P->L[0];
            s.off = 0;
            s.size = L.m;
            s.col = L.ell_col;
            s.val = vals(L);
            s.r = from_x ? (const void*)P->x : (const void*)P->p;
            s.x = into_tmp ? (void*)P->tmp : (void*)P->Ap;
            if(dot) s.ra = red(P, OP_PAP);
            return true;
        });
    }

    int vec(int kind, int op, std::function<void(Problem*, BSeg&)> set, double beta = 0.0)
    {
        return add(kind, RED_THREADS, true, 1, 0, beta, [&](int k, BSeg& s) {
            Problem* P = B.probs[k];
            s.size = P->L[0].m;
            set(P, s);
            if(op != OP_NONE) s.ra = red(P, op);
            return true;
        });
    }

    int true_residual(int op)
    {
        HPCGX_TRY(outer_spmv(false, true, true));
        HPCGX_TRY(vec(BK_WAXPBY, OP_NONE, [](Problem* P, BSeg& s) {
            s.a = P->b;
            s.b = P->tmp;
            s.w = P->r;
        }, -1.0));
        return vec(BK_DOT, op, [](Problem* P, BSeg& s) {
            s.a = P->r;
            s.b = P->r;
        });
    }
};

static int plan(Batch& B)
{
    for(int mode = 0; mode < 2; ++mode)
    {
        const bool mixed = mode == CG_MIXED;
        if(mixed ? !B.probs[0]->has_mixed : !B.probs[0]->has_f64) continue;
        const int prec = mixed ? PREC_MIXED : PREC_F64;
        // preamble (CG.cpp:133-137): p = x; Ap = A p; r = b - Ap; ||r||
        Planner pre{B, &B.pre[mode], prec};
        HPCGX_TRY(pre.vec(BK_COPY, OP_NONE, [](Problem* P, BSeg& s) {
            s.a = P->x;
            s.w = P->p;
        }));
        HPCGX_TRY(pre.outer_spmv(false, false, false));
        HPCGX_TRY(pre.vec(BK_WAXPBY, OP_NONE, [](Problem* P, BSeg& s) {
            s.a = P->b;
            s.b = P->Ap;
            s.w = P->r;
        }, -1.0));
        HPCGX_TRY(pre.vec(BK_DOT, OP_NORMR0, [](Problem* P, BSeg& s) {
            s.a = P->r;
            s.b = P->r;
        }));
        // one iteration (CG.cpp:150-171)
        Planner it{B, &B.iter[mode], prec};
        HPCGX_TRY(it.mg(0));
        it.next_bf32 = mixed;   // r (f64) . z (f32 in the mixed mode)
        HPCGX_TRY(it.vec(BK_DOT, OP_RTZ, [mixed](Problem* P, BSeg& s) {
            s.a = P->r;
            s.b = mixed ? (const void*)P->z32 : (const void*)P->z64;
        }));
        it.next_bf32 = mixed;
        HPCGX_TRY(it.vec(BK_CG_P, OP_NONE, [mixed](Problem* P, BSeg& s) {
            s.a = mixed ? (const void*)P->z32 : (const void*)P->z64;
            s.w = P->p;
            s.ra.sc = P->sc;
        }));
        HPCGX_TRY(it.outer_spmv(true, false, false));
        HPCGX_TRY(it.vec(BK_CG_XR, OP_NORMR, [](Problem* P, BSeg& s) {
            s.a = P->p;
            s.b = P->Ap;
            s.x = P->x;
            s.w = P->r;
        }));
    }
    B.planned = true;
    return 0;
}

// ---------------------------------------------------------------------------

static void free_plans(Batch& B)
{
    for(int g = 0; g < 2; ++g)
    {
        if(B.gexec[g]) (void)hipGraphExecDestroy(B.gexec[g]);
        if(B.graph[g]) (void)hipGraphDestroy(B.graph[g]);
        B.gexec[g] = nullptr;
        B.graph[g] = nullptr;
        B.pre[g].clear();
        B.iter[g].clear();
    }
    for(BSeg* t : B.tables) (void)hipFree(t);
    B.tables.clear();
    B.planned = false;
}

int batch_create(Batch*& out, int nprob, const int* dims, int nlevels, bool with_f64, bool with_mixed, hipStream_t s)
{
    out = nullptr;
    if(nprob < 1)
    {
        set_error("a batch needs at least one problem");
        return -2;
    }
    Batch* B = new Batch;
    B->nlevels = nlevels;
    if(s)
        B->s = s;
    else
    {
        if(hipStreamCreateWithFlags(&B->s, hipStreamNonBlocking) != hipSuccess)
        {
            set_error("cannot create the batch stream");
            delete B;
            return -1;
        }
        B->own_stream = true;
    }
    for(int k = 0; k < nprob; ++k)
    {
        Problem* P = nullptr;
        int rc = problem_create(P, dims[3 * k], dims[3 * k + 1], dims[3 * k + 2], nlevels, with_f64, with_mixed, B->s);
        if(rc)
        {
            batch_destroy(B);
            return rc;
        }
        B->probs.push_back(P);
    }
    out = B;
    return 0;
}

void batch_destroy(Batch* B)
{
    if(!B) return;
    if(B->s) (void)hipStreamSynchronize(B->s);
    free_plans(*B);
    for(Problem* P : B->probs) problem_destroy(P);
    if(B->own_stream && B->s) (void)hipStreamDestroy(B->s);
    delete B;
}

int batch_set_block(Batch& B, int level, int block)
{
    if(level < 0 || level >= B.nlevels || (block != 64 && block != 128 && block != 256))
    {
        set_error("invalid level or block size");
        return -2;
    }
    B.block[level] = block;
    HPCGX_HIP(hipStreamSynchronize(B.s));
    free_plans(B);
    return 0;
}

int batch_set_rows(Batch& B, int level, int rows)
{
    if(level < 0 || level >= B.nlevels || (rows != 1 && rows != 2))
    {
        set_error("invalid level or rows per thread");
        return -2;
    }
    B.rows[level] = rows;
    HPCGX_HIP(hipStreamSynchronize(B.s));
    free_plans(B);
    return 0;
}

static int run_list(const std::vector<BLaunch>& v, hipStream_t s)
{
    for(const BLaunch& L : v) HPCGX_TRY(launch(L, s));
    return 0;
}

// algorithmic HBM bytes of one planned launch (the per-row model of
// iteration_bytes in solver.hip, over the launch's rows)
static double launch_bytes(const BLaunch& L)
{
    const bool f64 = L.prec == PREC_F64;
    const double vb = f64 ? 8.0 : 2.0, xb = f64 ? 8.0 : 4.0, rb = L.r_f64 ? 8.0 : xb;
    const double slot = 4.0 + vb, n = (double)L.nrows;
    const double zb = L.bf32 ? 4.0 : 8.0;
    switch(L.kind)
    {
    case BK_SPMV: return n * (W * slot + 16.0);
    case BK_FWD_ZERO: return n * (window_slots(L.win) * slot + 4.0 + vb + rb + xb);
    case BK_BWD_ZERO: return n * (window_slots(L.win) * slot + 4.0 + vb + 2.0 * xb + (L.any_t ? xb : 0.0));
    case BK_FWD_FULL:
        if(L.all_u)
        {
            int na = 0, nb = 0;
            window_pair_slots(L.win, L.win_tail, &na, &nb);
            return n * ((na + nb) * slot + 4.0 + vb + rb + 3.0 * xb);
        }
        return n * (W * slot + rb + xb + (L.any_t ? xb : 0.0));
    case BK_BWD_FULL: return n * (window_slots(L.win) * slot + 4.0 + vb + 2.0 * xb);
    case BK_RESTRICT: return n * (W * slot + rb + 8.0 + xb);
    case BK_PROLONG: return n * (8.0 + 3.0 * xb);
    case BK_DOT: return n * (8.0 + zb);
    case BK_CG_P: return n * (zb + 16.0);
    case BK_CG_XR: return n * 48.0;
    case BK_COPY: return n * 16.0;
    case BK_WAXPBY: return n * 24.0;
    }
    return 0.0;
}

// cg_begin's reset of every problem's scalars (tolerance 0: run max_iter)
static int reset_scalars(Batch& B, int max_iter, double tol)
{
    for(Problem* P : B.probs)
    {
        std::memset(&P->h_sc, 0, sizeof(P->h_sc));
        P->h_sc.tol = tol;
        P->h_sc.max_iter = max_iter;
        HPCGX_HIP(hipMemcpyAsync(P->sc, &P->h_sc, sizeof(CGScalars), hipMemcpyHostToDevice, B.s));
    }
    HPCGX_HIP(hipStreamSynchronize(B.s));
    return 0;
}

int batch_profile(Batch& B, int mode, int iters, double* kind_ms, double* kind_bytes, double* iter_ms)
{
    if(mode != CG_F64 && mode != CG_MIXED)
    {
        set_error("invalid CG mode");
        return -2;
    }
    if(!B.planned) HPCGX_TRY(plan(B));
    const std::vector<BLaunch>& it = B.iter[mode];
    if(it.empty())
    {
        set_error("the batch was built without the store this mode needs");
        return -2;
    }
    if(iters < 1 || iters > B.probs[0]->hist_cap)
    {
        set_error("iteration count out of range");
        return -2;
    }
    HPCGX_TRY(reset_scalars(B, iters, 0.0));
    HPCGX_TRY(run_list(B.pre[mode], B.s));
    const size_t nl = it.size();
    std::vector<hipEvent_t> ev((size_t)iters * (nl + 1), nullptr);
    int rc = 0;
    for(auto& e : ev)
        if(hipEventCreate(&e) != hipSuccess)
        {
            set_error("hipEventCreate failed");
            rc = -1;
            break;
        }
    for(int i = 0; i < iters && !rc; ++i)
    {
        hipEvent_t* E = ev.data() + (size_t)i * (nl + 1);
        if(hipEventRecord(E[0], B.s) != hipSuccess) rc = -1;
        for(size_t j = 0; j < nl && !rc; ++j)
        {
            rc = launch(it[j], B.s);
            if(!rc && hipEventRecord(E[j + 1], B.s) != hipSuccess) rc = -1;
        }
    }
    if(!rc && hipStreamSynchronize(B.s) != hipSuccess)
    {
        set_error("stream synchronisation failed in a profiled batched run");
        rc = -1;
    }
    for(int k = 0; k < BK_COUNT; ++k)
    {
        kind_ms[k] = 0.0;
        kind_bytes[k] = 0.0;
    }
    double total = 0.0;
    if(!rc)
    {
        for(size_t j = 0; j < nl; ++j) kind_bytes[it[j].kind] += launch_bytes(it[j]);
        for(int i = 0; i < iters; ++i)
        {
            hipEvent_t* E = ev.data() + (size_t)i * (nl + 1);
            for(size_t j = 0; j < nl; ++j)
            {
                float ms = 0.f;
                if(hipEventElapsedTime(&ms, E[j], E[j + 1]) == hipSuccess) kind_ms[it[j].kind] += ms;
            }
            float ms = 0.f;
            if(hipEventElapsedTime(&ms, E[0], E[nl]) == hipSuccess) total += ms;
        }
    }
    for(auto& e : ev)
        if(e) (void)hipEventDestroy(e);
    if(iter_ms) *iter_ms = total / iters;
    return rc;
}

int batch_cg(Batch& B, int mode, int max_iter, double tol, bool use_graph, int* niters, double* normr, double* normr0)
{
    if(mode != CG_F64 && mode != CG_MIXED)
    {
        set_error("invalid CG mode");
        return -2;
    }
    if(!B.planned) HPCGX_TRY(plan(B));
    if(B.iter[mode].empty())
    {
        set_error("the batch was built without the store this mode needs");
        return -2;
    }
    if(max_iter < 0 || max_iter > B.probs[0]->hist_cap)
    {
        set_error("max_iter out of range");
        return -2;
    }
    // per-problem scalars (cg_begin's reset)
    HPCGX_TRY(reset_scalars(B, max_iter, tol));
    HPCGX_TRY(run_list(B.pre[mode], B.s));
    if(use_graph && !B.gexec[mode])
    {
        HPCGX_HIP(hipStreamBeginCapture(B.s, hipStreamCaptureModeThreadLocal));
        const int rc = run_list(B.iter[mode], B.s);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(B.s, &g);
        if(rc || e != hipSuccess)
        {
            if(g) (void)hipGraphDestroy(g);
            if(rc) return rc;
            set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
            return -1;
        }
        e = hipGraphInstantiate(&B.gexec[mode], g, nullptr, nullptr, 0);
        if(e != hipSuccess)
        {
            (void)hipGraphDestroy(g);
            B.gexec[mode] = nullptr;
            set_error(std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
            return -1;
        }
        B.graph[mode] = g;
    }
    const int check = 10;
    std::vector<CGScalars> h(B.probs.size());
    for(int it = 0; it < max_iter;)
    {
        const int n = std::min(check, max_iter - it);
        for(int i = 0; i < n; ++i)
        {
            if(use_graph)
                HPCGX_HIP(hipGraphLaunch(B.gexec[mode], B.s));
            else
                HPCGX_TRY(run_list(B.iter[mode], B.s));
        }
        it += n;
        if(tol <= 0.0 || it >= max_iter) continue;
        bool all = true;
        for(size_t k = 0; k < B.probs.size(); ++k)
            HPCGX_HIP(hipMemcpyAsync(&h[k], B.probs[k]->sc, sizeof(CGScalars), hipMemcpyDeviceToHost, B.s));
        HPCGX_HIP(hipStreamSynchronize(B.s));
        for(auto& v : h) all = all && v.done;
        if(all) break;
    }
    for(size_t k = 0; k < B.probs.size(); ++k)
        HPCGX_HIP(hipMemcpyAsync(&h[k], B.probs[k]->sc, sizeof(CGScalars), hipMemcpyDeviceToHost, B.s));
    HPCGX_HIP(hipStreamSynchronize(B.s));
    for(size_t k = 0; k < B.probs.size(); ++k)
    {
        if(niters) niters[k] = h[k].iter;
        if(normr) normr[k] = h[k].normr;
        if(normr0) normr0[k] = h[k].normr0;
    }
    return 0;
}

}  // namespace hpcgx
