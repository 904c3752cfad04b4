// kernels.hpp -- device helpers, the deterministic block reduction and the
// per-row bodies of the ELL operators, shared by the single-problem kernels
// (ops.hip) and the batched kernels (batch.hip).  Include from .hip files only.
#pragma once
#include "common.hpp"
#include "ops.hpp"

#include <type_traits>

namespace hpcgx
{

// ---------------------------------------------------------------------------
// scalar helpers
__device__ __forceinline__ double fmaT(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fmaT(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

template <typename T>
__device__ __forceinline__ T cvt(double v) { return (T)v; }
template <typename T>
__device__ __forceinline__ T cvt(float v) { return (T)v; }
template <typename T>
__device__ __forceinline__ T cvt(__half v) { return (T)__half2float(v); }

__device__ __forceinline__ bool gated(const int* gate) { return gate != nullptr && *gate != 0; }

// ---------------------------------------------------------------------------
// deterministic two-phase reductions (block partials, then one fold block);
// apply_op is what the fold does with the total.
__device__ inline void apply_op(int op, double v, CGScalars* sc, double* hist, int hist_cap, double* out)
{
    switch(op)
    {
    case OP_STORE: *out = v; break;
    case OP_NORMR0:
    {
        const double nr = sqrt(v);
        sc->normr = nr;
        sc->normr0 = nr;
        sc->iter = 0;
        if(hist && hist_cap >= 0) hist[0] = nr;
        sc->done = (!(nr / nr > sc->tol) || sc->max_iter <= 0) ? 1 : 0;
        break;
    }
    case OP_RTZ:
        sc->oldrtz = sc->rtz;
        sc->rtz = v;
        sc->beta = (sc->iter > 0) ? v / sc->oldrtz : 0.0;
        break;
    case OP_PAP:
        sc->pAp = v;
        sc->alpha = sc->rtz / v;
        break;
    case OP_NORMR:
    {
        const double nr = sqrt(v);
        sc->normr = nr;
        const int it = sc->iter + 1;
        sc->iter = it;
        if(hist && it <= hist_cap) hist[it] = nr;
        sc->done = (!(nr / sc->normr0 > sc->tol) || it >= sc->max_iter) ? 1 : 0;
        break;
    }
    case OP_NORMR_SET:
    {
        const double nr = sqrt(v);
        sc->normr = nr;
        const int it = sc->iter;
        if(hist && it <= hist_cap) hist[it] = nr;
        sc->done = (!(nr / sc->normr0 > sc->tol) || it >= sc->max_iter) ? 1 : 0;
        break;
    }
    default: break;
    }
}

// Phase 1: the block sum of v goes to partial[slot].
template <int BS>
__device__ __forceinline__ void block_partial(double v, double* __restrict__ partial, int64_t slot)
{
    __shared__ double lds[BS / 64];
    const double bsum = block_sum<BS>(v, lds);
    if(threadIdx.x == 0) partial[slot] = bsum;
}

// Phase 2, run by one block of BS threads in the kernel launched after the
// producer on the same stream (the kernel boundary orders the two): thread j
// sums partials j, j+BS, ... in four interleaved chains, the block sums the
// threads, thread 0 applies `op`.  The grouping depends only on n.
template <int BS>
__device__ __forceinline__ void fold_partials(const double* __restrict__ partial, int64_t n, const RedArgs& ra)
{
    __shared__ double lds[BS / 64];
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int64_t i = threadIdx.x;
    for(; i + 3 * BS < n; i += 4 * BS)
    {
        a0 += partial[i];
        a1 += partial[i + BS];
        a2 += partial[i + 2 * BS];
        a3 += partial[i + 3 * BS];
    }
    for(; i < n; i += BS) a0 += partial[i];
    const double s = block_sum<BS>((a0 + a1) + (a2 + a3), lds);
    if(threadIdx.x == 0) apply_op(ra.op, s, ra.sc, ra.hist, ra.hist_cap, ra.out);
}

// ---------------------------------------------------------------------------
// Vector loads of R consecutive rows of one ELL slot.  A thread owns rows
// [row, row+R) (row % R == 0, m % R == 0), so a wave issues one contiguous
// 64*R-element segment per slot: dwordx2/x4 loads instead of R dword loads.
template <typename T, int R>
struct Ext;
#define HPCGX_EXT(T, R) \
    template <>         \
    struct Ext<T, R>    \
    {                   \
        typedef T type __attribute__((ext_vector_type(R))); \
    };
HPCGX_EXT(int32_t, 2)
HPCGX_EXT(int32_t, 4)
HPCGX_EXT(double, 2)
HPCGX_EXT(double, 4)
HPCGX_EXT(float, 2)
HPCGX_EXT(float, 4)
HPCGX_EXT(uint16_t, 2)
HPCGX_EXT(uint16_t, 4)
#undef HPCGX_EXT

template <int R, typename T>
__device__ __forceinline__ void ld_nt(const T* p, T (&o)[R])
{
    if constexpr(R == 1)
        o[0] = ntload(p);
    else
    {
        typedef typename Ext<T, R>::type V;
        const V w = __builtin_nontemporal_load(reinterpret_cast<const V*>(p));
#pragma unroll
        for(int k = 0; k < R; ++k) o[k] = w[k];
    }
}

template <int R>
__device__ __forceinline__ void ld_nt(const __half* p, __half (&o)[R])
{
    if constexpr(R == 1)
        o[0] = ntload(p);
    else
    {
        typedef typename Ext<uint16_t, R>::type V;
        const V w = __builtin_nontemporal_load(reinterpret_cast<const V*>(p));
#pragma unroll
        for(int k = 0; k < R; ++k) o[k] = __ushort_as_half(w[k]);
    }
}

template <int R, typename T>
__device__ __forceinline__ void ld(const T* p, T (&o)[R])
{
    if constexpr(R == 1)
        o[0] = *p;
    else
    {
        typedef typename Ext<T, R>::type V;
        const V w = *reinterpret_cast<const V*>(p);
#pragma unroll
        for(int k = 0; k < R; ++k) o[k] = w[k];
    }
}

template <int R, typename T>
__device__ __forceinline__ void st(T* p, const T (&v)[R])
{
    if constexpr(R == 1)
        *p = v[0];
    else
    {
        typedef typename Ext<T, R>::type V;
        V w;
#pragma unroll
        for(int k = 0; k < R; ++k) w[k] = v[k];
        *reinterpret_cast<V*>(p) = w;
    }
}

// r may be f64 while x is f32 (mixed level 0): load in RT, convert.
template <int R, typename XT, typename RT>
__device__ __forceinline__ void ld_cvt(const RT* p, XT (&o)[R])
{
    RT t[R];
    ld<R>(p, t);
#pragma unroll
    for(int k = 0; k < R; ++k) o[k] = (XT)t[k];
}

template <int R, typename XT, typename YT>
__device__ __forceinline__ void st_cvt(YT* p, const XT (&v)[R])
{
    YT t[R];
#pragma unroll
    for(int k = 0; k < R; ++k) t[k] = (YT)v[k];
    st<R>(p, t);
}

// ---------------------------------------------------------------------------
// Row tiles of a whole-level operator: tile t -> (slab j = t / ncolors,
// colour c = t % ncolors), rows offsets[c] + j*BS*R + [0, BS*R) clipped to the
// colour.  Consecutive tiles cover the same spatial slab in every colour, and
// consecutive blocks of an XCD take consecutive runs of `tpb` tiles, so the
// x gathers of the rows in flight stay inside a compact slab.
struct Tiles
{
    int ncolors;
    int jtiles;   // max over colours of ceil(size / (BS*R))
    int tpb;      // tiles per block
    int off[MAX_COLORS];
    int size[MAX_COLORS];
};

// ---------------------------------------------------------------------------
// Row bodies shared by the single-problem kernels (ops.hip) and the batched
// kernels (batch.hip).  A thread owns rows [row, row+R) (R-aligned).

// y = A x on R rows; returns the row sums in s[].
template <typename VT, typename XT, int R>
__device__ __forceinline__ void row_spmv(int64_t m, int64_t row, const int32_t* __restrict__ col,
                                         const VT* __restrict__ val, const XT* __restrict__ x, XT (&s)[R])
{
    int32_t cc[W][R];
    VT av[W][R];
#pragma unroll
    for(int p = 0; p < W; ++p) ld_nt<R>(col + (int64_t)p * m + row, cc[p]);
#pragma unroll
    for(int p = 0; p < W; ++p) ld_nt<R>(val + (int64_t)p * m + row, av[p]);
    XT xv[W][R];
#pragma unroll
    for(int p = 0; p < W; ++p)
#pragma unroll
        for(int k = 0; k < R; ++k) xv[p][k] = (cc[p][k] >= 0) ? x[cc[p][k]] : (XT)0;
#pragma unroll
    for(int k = 0; k < R; ++k) s[k] = 0;
#pragma unroll
    for(int p = 0; p < W; ++p)
#pragma unroll
        for(int k = 0; k < R; ++k)
            if(cc[p][k] >= 0) s[k] = fmaT(cvt<XT>(av[p][k]), xv[p][k], s[k]);
}

// Row descriptor, one int32 per row built with the ELL (slots hold the
// columns in ascending multicolour order, padding last): bits 0-7 the slot of
// the diagonal d, bits 8-15 the first slot whose column lies in the last
// colour tl (the end when there is none), bits 16-23 the number of valid
// slots e.  So [0, d) are the lower entries, (d, tl) the upper entries of the
// colours before the last and [tl, e) the entries of the last colour.
__device__ __forceinline__ int desc_diag(int32_t v) { return v & 255; }
__device__ __forceinline__ int desc_tail(int32_t v) { return (v >> 8) & 255; }
__device__ __forceinline__ int desc_end(int32_t v) { return (v >> 16) & 255; }

// Loads slot p of the thread's R rows where in(p, k) holds for some k, and
// marks the slots of row k outside in(p, k) with column -1.
template <typename VT, int R, typename In>
__device__ __forceinline__ void load_slots(int64_t m, int64_t row, const int32_t* __restrict__ col,
                                           const VT* __restrict__ val, int32_t (&cc)[W][R], VT (&av)[W][R], In in)
{
#pragma unroll
    for(int p = 0; p < W; ++p)
    {
        bool any = false;
#pragma unroll
        for(int k = 0; k < R; ++k) any = any || in(p, k);
        if(any)
        {
            ld_nt<R>(col + (int64_t)p * m + row, cc[p]);
            ld_nt<R>(val + (int64_t)p * m + row, av[p]);
        }
#pragma unroll
        for(int k = 0; k < R; ++k)
            if(!in(p, k)) cc[p][k] = -1;
    }
}

template <typename XT, int R>
__device__ __forceinline__ void gather(const int32_t (&cc)[W][R], const XT* __restrict__ x, XT (&xv)[W][R])
{
#pragma unroll
    for(int p = 0; p < W; ++p)
#pragma unroll
        for(int k = 0; k < R; ++k) xv[p][k] = (cc[p][k] >= 0) ? x[cc[p][k]] : (XT)0;
}

// s[k] -= sum over the marked slots of row k, ascending
template <typename VT, typename XT, int R>
__device__ __forceinline__ void sub_slots(const int32_t (&cc)[W][R], const VT (&av)[W][R], const XT (&xv)[W][R],
                                          XT (&s)[R])
{
#pragma unroll
    for(int p = 0; p < W; ++p)
#pragma unroll
        for(int k = 0; k < R; ++k)
            if(cc[p][k] >= 0) s[k] = fmaT(-cvt<XT>(av[p][k]), xv[p][k], s[k]);
}

template <typename VT, typename XT, int R>
__device__ __forceinline__ void diag_vals(int64_t m, int64_t row, const VT* __restrict__ val, const int32_t (&ds)[R],
                                          XT (&d)[R])
{
#pragma unroll
    for(int k = 0; k < R; ++k) d[k] = cvt<XT>(val[(int64_t)desc_diag(ds[k]) * m + row + k]);
}

// Zero-initial-guess forward sweep: only entries left of the diagonal
// (earlier colours) are non-zero in x.
template <typename VT, typename XT, typename RT, int R>
__device__ __forceinline__ void row_fwd_zero(int64_t m, int64_t row, const int32_t* __restrict__ col,
                                             const VT* __restrict__ val, const int32_t* __restrict__ desc,
                                             const RT* __restrict__ r, XT* __restrict__ x)
{
    int32_t ds[R];
    ld<R>(desc + row, ds);
    int32_t cc[W][R];
    VT av[W][R];
    load_slots<VT, R>(m, row, col, val, cc, av, [&](int p, int k) { return p < desc_diag(ds[k]); });
    XT d[R], s[R];
    diag_vals<VT, XT, R>(m, row, val, ds, d);
    ld_cvt<R>(r + row, s);
    XT xv[W][R];
    gather<XT, R>(cc, x, xv);
    sub_slots<VT, XT, R>(cc, av, xv, s);
#pragma unroll
    for(int k = 0; k < R; ++k) s[k] = s[k] / d[k];
    st<R>(x + row, s);
}

// Zero-initial-guess backward sweep: x = (x*d - U x) / d.  With STORE_U it
// also keeps u = sum over (d, tl) of a*x -- the upper entries outside the last
// colour, whose x the coarse-grid correction leaves alone -- for the
// post-smoother's forward sweep (row_fwd_full_u).
template <typename VT, typename XT, bool STORE_U, int R>
__device__ __forceinline__ void row_bwd_zero(int64_t m, int64_t row, const int32_t* __restrict__ col,
                                             const VT* __restrict__ val, const int32_t* __restrict__ desc,
                                             XT* __restrict__ x, XT* __restrict__ u)
{
    int32_t ds[R];
    ld<R>(desc + row, ds);
    int32_t cc[W][R];
    VT av[W][R];
    load_slots<VT, R>(m, row, col, val, cc, av,
                      [&](int p, int k) { return p > desc_diag(ds[k]) && p < desc_end(ds[k]); });
    XT d[R], s[R];
    diag_vals<VT, XT, R>(m, row, val, ds, d);
    ld<R>(x + row, s);
#pragma unroll
    for(int k = 0; k < R; ++k) s[k] = s[k] * d[k];
    XT xv[W][R];
    gather<XT, R>(cc, x, xv);
    sub_slots<VT, XT, R>(cc, av, xv, s);
#pragma unroll
    for(int k = 0; k < R; ++k) s[k] = s[k] / d[k];
    st<R>(x + row, s);
    if(STORE_U)
    {
        XT uu[R];
#pragma unroll
        for(int k = 0; k < R; ++k) uu[k] = 0;
#pragma unroll
        for(int p = 0; p < W; ++p)
#pragma unroll
            for(int k = 0; k < R; ++k)
                if(cc[p][k] >= 0 && p < desc_tail(ds[k])) uu[k] = fmaT(cvt<XT>(av[p][k]), xv[p][k], uu[k]);
        st<R>(u + row, uu);
    }
}

// Full forward sweep over the whole row in ascending column order; keeps
// t = r - L x (the sum up to the diagonal) for the backward sweep, whose lower
// neighbours do not change in between.
template <typename VT, typename XT, typename RT, bool STORE_T, int R>
__device__ __forceinline__ void row_fwd_full(int64_t m, int64_t row, const int32_t* __restrict__ col,
                                             const VT* __restrict__ val, const RT* __restrict__ r,
                                             XT* __restrict__ x, XT* __restrict__ t)
{
    int32_t cc[W][R];
    VT av[W][R];
#pragma unroll
    for(int p = 0; p < W; ++p) ld_nt<R>(col + (int64_t)p * m + row, cc[p]);
#pragma unroll
    for(int p = 0; p < W; ++p) ld_nt<R>(val + (int64_t)p * m + row, av[p]);
    XT xv[W][R];
#pragma unroll
    for(int p = 0; p < W; ++p)
#pragma unroll
        for(int k = 0; k < R; ++k)
            xv[p][k] = (cc[p][k] >= 0 && (int64_t)cc[p][k] != row + k) ? x[cc[p][k]] : (XT)0;
    XT s[R], tt[R], d[R];
    ld_cvt<R>(r + row, s);
#pragma unroll
    for(int k = 0; k < R; ++k)
    {
        tt[k] = 0;
        d[k] = 1;
    }
#pragma unroll
    for(int p = 0; p < W; ++p)
#pragma unroll
        for(int k = 0; k < R; ++k)
        {
            if(cc[p][k] >= 0)
            {
                if((int64_t)cc[p][k] == row + k)
                {
                    tt[k] = s[k];
                    d[k] = cvt<XT>(av[p][k]);
                }
                else
                {
                    s[k] = fmaT(-cvt<XT>(av[p][k]), xv[p][k], s[k]);
                }
            }
        }
#pragma unroll
    for(int k = 0; k < R; ++k) s[k] = s[k] / d[k];
    st<R>(x + row, s);
    if(STORE_T) st<R>(t + row, tt);
}

// The post-smoother's forward sweep for a row of a colour before the last,
// when the coarse-grid correction changed only rows of the last colour:
// x = (r - L x - u - T x) / d with u = t[row] from row_bwd_zero<STORE_U> (the
// upper entries outside the last colour at their unchanged x) and T the
// last-colour entries [tl, e) at their corrected x.  Reads the lower, the
// diagonal and the last-colour slots only; t[row] becomes r - L x.
template <typename VT, typename XT, typename RT, int R>
__device__ __forceinline__ void row_fwd_full_u(int64_t m, int64_t row, const int32_t* __restrict__ col,
                                               const VT* __restrict__ val, const int32_t* __restrict__ desc,
                                               const RT* __restrict__ r, XT* __restrict__ x, XT* __restrict__ t)
{
    int32_t ds[R];
    ld<R>(desc + row, ds);
    int32_t cc[W][R];
    VT av[W][R];
    load_slots<VT, R>(m, row, col, val, cc, av, [&](int p, int k) {
        return p < desc_diag(ds[k]) || (p >= desc_tail(ds[k]) && p < desc_end(ds[k]));
    });
    XT d[R], s[R], uu[R];
    diag_vals<VT, XT, R>(m, row, val, ds, d);
    ld_cvt<R>(r + row, s);
    ld<R>(t + row, uu);
    XT xv[W][R];
    gather<XT, R>(cc, x, xv);
#pragma unroll
    for(int p = 0; p < W; ++p)
#pragma unroll
        for(int k = 0; k < R; ++k)
            if(cc[p][k] >= 0 && p < desc_diag(ds[k])) s[k] = fmaT(-cvt<XT>(av[p][k]), xv[p][k], s[k]);
    XT tt[R];
#pragma unroll
    for(int k = 0; k < R; ++k)
    {
        tt[k] = s[k];
        s[k] = s[k] - uu[k];
    }
#pragma unroll
    for(int p = 0; p < W; ++p)
#pragma unroll
        for(int k = 0; k < R; ++k)
            if(cc[p][k] >= 0 && p > desc_diag(ds[k])) s[k] = fmaT(-cvt<XT>(av[p][k]), xv[p][k], s[k]);
#pragma unroll
    for(int k = 0; k < R; ++k) s[k] = s[k] / d[k];
    st<R>(x + row, s);
    st<R>(t + row, tt);
}

// Full backward sweep from t: x = (t - U x) / d.
template <typename VT, typename XT, int R>
__device__ __forceinline__ void row_bwd_full(int64_t m, int64_t row, const int32_t* __restrict__ col,
                                             const VT* __restrict__ val, const int32_t* __restrict__ desc,
                                             const XT* __restrict__ t, XT* __restrict__ x)
{
    int32_t ds[R];
    ld<R>(desc + row, ds);
    int32_t cc[W][R];
    VT av[W][R];
    load_slots<VT, R>(m, row, col, val, cc, av,
                      [&](int p, int k) { return p > desc_diag(ds[k]) && p < desc_end(ds[k]); });
    XT d[R], s[R];
    diag_vals<VT, XT, R>(m, row, val, ds, d);
    ld<R>(t + row, s);
    XT xv[W][R];
    gather<XT, R>(cc, x, xv);
    sub_slots<VT, XT, R>(cc, av, xv, s);
#pragma unroll
    for(int k = 0; k < R; ++k) s[k] = s[k] / d[k];
    st<R>(x + row, s);
}

// Fused fine residual + injection for coarse natural row k:
// rc[f2c_coarse[k]] = r[f] - sum_p a[f,p] x[col], f = f2c_fine[k].
template <typename VT, typename XT, typename RT, typename CT>
__device__ __forceinline__ void row_restrict(int64_t m, int64_t k, const int32_t* __restrict__ col,
                                             const VT* __restrict__ val, const int32_t* __restrict__ f2c_fine,
                                             const int32_t* __restrict__ f2c_coarse, const RT* __restrict__ r,
                                             const XT* __restrict__ x, CT* __restrict__ rc)
{
    const int64_t row = f2c_fine[k];
    int32_t cc[W][1];
    VT av[W][1];
#pragma unroll
    for(int p = 0; p < W; ++p) ld_nt<1>(col + (int64_t)p * m + row, cc[p]);
#pragma unroll
    for(int p = 0; p < W; ++p) ld_nt<1>(val + (int64_t)p * m + row, av[p]);
    XT xv[W];
#pragma unroll
    for(int p = 0; p < W; ++p) xv[p] = (cc[p][0] >= 0) ? x[cc[p][0]] : (XT)0;
    XT s = cvt<XT>(r[row]);
#pragma unroll
    for(int p = 0; p < W; ++p)
        if(cc[p][0] >= 0) s = fmaT(-cvt<XT>(av[p][0]), xv[p], s);
    rc[f2c_coarse[k]] = (CT)s;
}

}  // namespace hpcgx
