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
        sc->xpend = 0;   // the p update before this SpMV applied the pending x update
        break;
    case OP_NORMR:
    {
        const double nr = sqrt(v);
        sc->normr = nr;
        sc->xpend = 1;   // x += alpha*p waits for the next p update (or the end)
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
        sc->xpend = 0;   // the true residual was taken from an updated x
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

template <int R>
__device__ __forceinline__ void ld(const __half* p, __half (&o)[R])
{
    if constexpr(R == 1)
        o[0] = *p;
    else
    {
        typedef typename Ext<uint16_t, R>::type V;
        const V w = *reinterpret_cast<const V*>(p);
#pragma unroll
        for(int k = 0; k < R; ++k) o[k] = __ushort_as_half(w[k]);
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
// the blocks in flight on an XCD work on neighbouring tiles, so the x gathers
// of the rows in flight stay inside a compact slab that the XCD's L2 holds.
// Blocks are numbered XCD-contiguously (xcd_block).  With tpb tiles per block,
// consecutive blocks form groups of gx (about what an XCD keeps resident);
// a group owns gx*tpb consecutive tiles and its blocks step through them side
// by side (member l takes tiles base + l, + gx, + 2 gx, ...).
struct Tiles
{
    int ncolors;
    int jtiles;   // max over colours of ceil(size / (BS*R))
    int tpb;      // tiles per block
    int gx;       // blocks per group
    int off[MAX_COLORS];
    int size[MAX_COLORS];
};

// ---------------------------------------------------------------------------
// Row bodies shared by the single-problem kernels (ops.hip) and the batched
// kernels (batch.hip).  A thread owns rows [row, row+R) (R-aligned).

// y = A x on R rows; returns the row sums in s[].
template <typename VT, typename XT, int R>
__device__ __forceinline__ void row_spmv(int64_t row, const int32_t* __restrict__ col,
                                         const VT* __restrict__ val, const XT* __restrict__ x, XT (&s)[R])
{
    int32_t cc[W][R];
    VT av[W][R];
#pragma unroll
    for(int p = 0; p < W; ++p) ld_nt<R>(col + ell_at(p, row), cc[p]);
#pragma unroll
    for(int p = 0; p < W; ++p) ld_nt<R>(val + ell_at(p, row), av[p]);
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

// ELL layout of the sweeps ("split" padding).  The entries of a row keep
// the ascending multicolour column order of the reference's ELL, but the
// padding sits between the diagonal and the upper entries instead of at the
// tail: slots [0, d) hold the lower entries, slot d the diagonal, slots
// [us, W) the upper entries, right-aligned, and the last-colour entries (the
// largest columns) are their tail [tl, W).  Every window a sweep reads then
// starts at a slot that does not depend on the row -- 0 for the lower
// entries, W - NS for the upper or last-colour ones -- so its loads are issued
// with the row's descriptor load instead of after it.
//
// Row descriptor, one int32 per row: bits 0-7 the diagonal slot d, bits 8-15
// the first slot of an off-diagonal last-colour column tl, bits 16-23 the
// first upper slot us (W when there is none).
__device__ __forceinline__ int desc_diag(int32_t v) { return v & 255; }
__device__ __forceinline__ int desc_tail(int32_t v) { return (v >> 8) & 255; }
__device__ __forceinline__ int desc_upper(int32_t v) { return (v >> 16) & 255; }

// Slot windows.  A sweep reads NS consecutive slots from a fixed start: 0
// for the lower entries, W - NS for the upper or last-colour ones.  NS is a
// compile-time bound the host picks per colour (Level::win, rounded up to a
// bucket; the buckets are the windows the parity colourings produce, so no
// slot is read in vain there), so light colours keep few registers live.
// The loads are unconditional -- no branch between them, so they are all in
// flight at once -- and which loaded slots a row uses is decided afterwards
// from its descriptor.  NS may be 0 (a colour with no lower entries); arrays
// then keep one unused element.
constexpr int slot_dim(int n) { return n > 0 ? n : 1; }

template <int NS, typename VT, int R>
__device__ __forceinline__ void load_window(int64_t row, const int32_t* __restrict__ col,
                                            const VT* __restrict__ val, int start,
                                            int32_t (&cc)[slot_dim(NS)][R], VT (&av)[slot_dim(NS)][R])
{
#pragma unroll
    for(int q = 0; q < NS; ++q)
    {
        ld_nt<R>(col + ell_at(start + q, row), cc[q]);
        ld_nt<R>(val + ell_at(start + q, row), av[q]);
    }
}

// Neighbour values of every loaded slot with a column (a slot the row does
// not use reads a value that is then ignored: its own diagonal, or a
// neighbour outside the window's kind -- never a row this launch writes).
template <int NS, typename XT, int R>
__device__ __forceinline__ void gather(const int32_t (&cc)[slot_dim(NS)][R], const XT* __restrict__ x,
                                       XT (&xv)[slot_dim(NS)][R])
{
#pragma unroll
    for(int q = 0; q < NS; ++q)
#pragma unroll
        for(int k = 0; k < R; ++k) xv[q][k] = (cc[q][k] >= 0) ? x[cc[q][k]] : (XT)0;
}

// s -= a x over the window's slots in [lo, hi) of each row, ascending.
template <int NS, typename VT, typename XT, int R>
__device__ __forceinline__ void sub_window(int start, const int (&lo)[R], const int (&hi)[R],
                                           const int32_t (&cc)[slot_dim(NS)][R],
                                           const VT (&av)[slot_dim(NS)][R], const XT (&xv)[slot_dim(NS)][R],
                                           XT (&s)[R])
{
#pragma unroll
    for(int q = 0; q < NS; ++q)
#pragma unroll
        for(int k = 0; k < R; ++k)
            if(start + q >= lo[k] && start + q < hi[k] && cc[q][k] >= 0)
                s[k] = fmaT(-cvt<XT>(av[q][k]), xv[q][k], s[k]);
}

// Zero-initial-guess forward sweep: only entries left of the diagonal
// (earlier colours) are non-zero in x.  Window [0, NS), NS >= the colour's
// largest d; the diagonal comes from its own array (dg).
template <int NS, typename VT, typename XT, typename RT, int R>
__device__ __forceinline__ void row_fwd_zero(int64_t row, const int32_t* __restrict__ col,
                                             const VT* __restrict__ val, const VT* __restrict__ dg,
                                             const int32_t* __restrict__ desc, const RT* __restrict__ r,
                                             XT* __restrict__ x)
{
    int32_t ds[R];
    ld<R>(desc + row, ds);
    int32_t cc[slot_dim(NS)][R];
    VT av[slot_dim(NS)][R];
    load_window<NS, VT, R>(row, col, val, 0, cc, av);
    XT d[R], s[R];
    ld_cvt<R>(dg + row, d);
    ld_cvt<R>(r + row, s);
    XT xv[slot_dim(NS)][R];
    gather<NS, XT, R>(cc, x, xv);
    int lo[R], hi[R];
#pragma unroll
    for(int k = 0; k < R; ++k)
    {
        lo[k] = 0;
        hi[k] = desc_diag(ds[k]);
    }
    sub_window<NS, VT, XT, R>(0, lo, hi, cc, av, xv, s);
#pragma unroll
    for(int k = 0; k < R; ++k) s[k] = s[k] / d[k];
    st<R>(x + row, s);
}

// Zero-initial-guess backward sweep: x = (x*d - U x) / d over the upper
// entries, window [W - NS, W), NS >= the colour's largest upper count.  With
// STORE_U it also keeps u = the sum over [us, tl) of a*x -- the upper entries
// outside the last colour, whose x the coarse-grid correction leaves alone --
// for the post-smoother's forward sweep (row_fwd_full_u); u is read off the
// running sum (x*d minus its value on reaching slot tl), so it costs no second
// accumulation.
template <int NS, typename VT, typename XT, bool STORE_U, int R>
__device__ __forceinline__ void row_bwd_zero(int64_t row, const int32_t* __restrict__ col,
                                             const VT* __restrict__ val, const VT* __restrict__ dg,
                                             const int32_t* __restrict__ desc, XT* __restrict__ x,
                                             XT* __restrict__ u)
{
    // u = x*d - snap must use the rounded product that starts the running
    // sum: left to the compiler, the product may or may not be contracted
    // into the subtraction depending on the instruction schedule, and then
    // two launch shapes of the same operator round u differently.
#pragma clang fp contract(off)
    int32_t ds[R];
    ld<R>(desc + row, ds);
    const int start = W - NS;
    int32_t cc[slot_dim(NS)][R];
    VT av[slot_dim(NS)][R];
    load_window<NS, VT, R>(row, col, val, start, cc, av);
    XT d[R], s[R];
    ld_cvt<R>(dg + row, d);
    ld<R>(x + row, s);
#pragma unroll
    for(int k = 0; k < R; ++k) s[k] = s[k] * d[k];
    XT xv[slot_dim(NS)][R];
    gather<NS, XT, R>(cc, x, xv);
    int lo[R], hi[R];
#pragma unroll
    for(int k = 0; k < R; ++k)
    {
        lo[k] = desc_upper(ds[k]);
        hi[k] = W;
    }
    if constexpr(STORE_U)
    {
        // u = x*d - (the running sum when it reaches the first last-colour
        // slot): one snapshot register per row instead of a second chain
        XT s0[R], snap[R];
        bool taken[R];
#pragma unroll
        for(int k = 0; k < R; ++k)
        {
            s0[k] = s[k];
            snap[k] = s[k];
            taken[k] = false;
        }
#pragma unroll
        for(int q = 0; q < NS; ++q)
#pragma unroll
            for(int k = 0; k < R; ++k)
            {
                if(!taken[k] && start + q >= desc_tail(ds[k]))
                {
                    snap[k] = s[k];
                    taken[k] = true;
                }
                if(start + q >= lo[k] && cc[q][k] >= 0) s[k] = fmaT(-cvt<XT>(av[q][k]), xv[q][k], s[k]);
            }
        XT uu[R];
#pragma unroll
        for(int k = 0; k < R; ++k) uu[k] = s0[k] - (taken[k] ? snap[k] : s[k]);
        st<R>(u + row, uu);
    }
    else
    {
        sub_window<NS, VT, XT, R>(start, lo, hi, cc, av, xv, s);
    }
#pragma unroll
    for(int k = 0; k < R; ++k) s[k] = s[k] / d[k];
    st<R>(x + row, s);
}

// Full forward sweep over the whole row in ascending column order; keeps
// t = r - L x (the sum up to the diagonal) for the backward sweep, whose lower
// neighbours do not change in between.
template <typename VT, typename XT, typename RT, bool STORE_T, int R>
__device__ __forceinline__ void row_fwd_full(int64_t row, const int32_t* __restrict__ col,
                                             const VT* __restrict__ val, const RT* __restrict__ r,
                                             XT* __restrict__ x, XT* __restrict__ t)
{
    int32_t cc[W][R];
    VT av[W][R];
#pragma unroll
    for(int p = 0; p < W; ++p) ld_nt<R>(col + ell_at(p, row), cc[p]);
#pragma unroll
    for(int p = 0; p < W; ++p) ld_nt<R>(val + ell_at(p, row), av[p]);
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
// last-colour entries [tl, W) at their corrected x.  Reads the lower, the
// diagonal and the last-colour slots only: windows [0, NA) and [W - NB, W);
// t[row] becomes r - L x.
template <int NA, int NB, typename VT, typename XT, typename RT, int R>
__device__ __forceinline__ void row_fwd_full_u(int64_t row, const int32_t* __restrict__ col,
                                               const VT* __restrict__ val,
                                               const VT* __restrict__ dg, const int32_t* __restrict__ desc,
                                               const RT* __restrict__ r, XT* __restrict__ x, XT* __restrict__ t)
{
    int32_t ds[R];
    ld<R>(desc + row, ds);
    const int bstart = W - NB;
    int32_t ca[slot_dim(NA)][R], cb[slot_dim(NB)][R];
    VT aa[slot_dim(NA)][R], ab[slot_dim(NB)][R];
    load_window<NA, VT, R>(row, col, val, 0, ca, aa);
    load_window<NB, VT, R>(row, col, val, bstart, cb, ab);
    XT d[R], s[R], uu[R];
    ld_cvt<R>(dg + row, d);
    ld_cvt<R>(r + row, s);
    ld<R>(t + row, uu);
    XT xa[slot_dim(NA)][R], xb[slot_dim(NB)][R];
    gather<NA, XT, R>(ca, x, xa);
    gather<NB, XT, R>(cb, x, xb);
    int alo[R], ahi[R], blo[R], bhi[R];
#pragma unroll
    for(int k = 0; k < R; ++k)
    {
        alo[k] = 0;
        ahi[k] = desc_diag(ds[k]);
        blo[k] = desc_tail(ds[k]);
        bhi[k] = W;
    }
    sub_window<NA, VT, XT, R>(0, alo, ahi, ca, aa, xa, s);
    XT tt[R];
#pragma unroll
    for(int k = 0; k < R; ++k)
    {
        tt[k] = s[k];
        s[k] = s[k] - uu[k];
    }
    sub_window<NB, VT, XT, R>(bstart, blo, bhi, cb, ab, xb, s);
#pragma unroll
    for(int k = 0; k < R; ++k) s[k] = s[k] / d[k];
    st<R>(x + row, s);
    st<R>(t + row, tt);
}

// Full backward sweep from t: x = (t - U x) / d over the upper entries,
// window [W - NS, W).
template <int NS, typename VT, typename XT, int R>
__device__ __forceinline__ void row_bwd_full(int64_t row, const int32_t* __restrict__ col,
                                             const VT* __restrict__ val, const VT* __restrict__ dg,
                                             const int32_t* __restrict__ desc, const XT* __restrict__ t,
                                             XT* __restrict__ x)
{
    int32_t ds[R];
    ld<R>(desc + row, ds);
    const int start = W - NS;
    int32_t cc[slot_dim(NS)][R];
    VT av[slot_dim(NS)][R];
    load_window<NS, VT, R>(row, col, val, start, cc, av);
    XT d[R], s[R];
    ld_cvt<R>(dg + row, d);
    ld<R>(t + row, s);
    XT xv[slot_dim(NS)][R];
    gather<NS, XT, R>(cc, x, xv);
    int lo[R], hi[R];
#pragma unroll
    for(int k = 0; k < R; ++k)
    {
        lo[k] = desc_upper(ds[k]);
        hi[k] = W;
    }
    sub_window<NS, VT, XT, R>(start, lo, hi, cc, av, xv, s);
#pragma unroll
    for(int k = 0; k < R; ++k) s[k] = s[k] / d[k];
    st<R>(x + row, s);
}

// Fused fine residual + injection for coarse natural row k:
// rc[f2c_coarse[k]] = r[f] - sum_p a[f,p] x[col], f = f2c_fine[k].
template <typename VT, typename XT, typename RT, typename CT>
__device__ __forceinline__ void row_restrict(int64_t k, const int32_t* __restrict__ col,
                                             const VT* __restrict__ val, const int32_t* __restrict__ f2c_fine,
                                             const int32_t* __restrict__ f2c_coarse, const RT* __restrict__ r,
                                             const XT* __restrict__ x, CT* __restrict__ rc)
{
    const int64_t row = f2c_fine[k];
    int32_t cc[W][1];
    VT av[W][1];
#pragma unroll
    for(int p = 0; p < W; ++p) ld_nt<1>(col + ell_at(p, row), cc[p]);
#pragma unroll
    for(int p = 0; p < W; ++p) ld_nt<1>(val + ell_at(p, row), av[p]);
    XT xv[W];
#pragma unroll
    for(int p = 0; p < W; ++p) xv[p] = (cc[p][0] >= 0) ? x[cc[p][0]] : (XT)0;
    XT s = cvt<XT>(r[row]);
#pragma unroll
    for(int p = 0; p < W; ++p)
        if(cc[p][0] >= 0) s = fmaT(-cvt<XT>(av[p][0]), xv[p], s);
    rc[f2c_coarse[k]] = (CT)s;
}

// ---------------------------------------------------------------------------
// host side: dispatch of the slot-window bounds (shared by ops.hip and
// batch.hip)

// Compile-time slot-window bounds.  The buckets are the window sizes of the
// 27-point operator's parity colourings (lower: 0 2 10 12 16 22 26, upper:
// 4 10 14 16 24 26 per colour), so those colours read no slot in vain and
// keep only the registers they need; any other window takes the next bucket
// up (at most 26: the diagonal always has a slot of its own).
template <typename F>
static void with_window(int w, F&& f)
{
    using std::integral_constant;
#define HPCGX_WIN(N)                          \
    if(w <= (N))                              \
    {                                         \
        f(integral_constant<int, N>{});       \
        return;                               \
    }
    HPCGX_WIN(0)
    HPCGX_WIN(2)
    HPCGX_WIN(4)
    HPCGX_WIN(10)
    HPCGX_WIN(12)
    HPCGX_WIN(14)
    HPCGX_WIN(16)
    HPCGX_WIN(22)
    HPCGX_WIN(24)
#undef HPCGX_WIN
    f(integral_constant<int, W - 1>{});
}

// (lower, last-colour) window pairs of the post-smoother's forward sweep, as
// they occur per colour of the parity colourings; anything else takes the
// first listed pair covering it, or (26, 26)
template <typename F>
static void with_window_pair(int wl, int wt, F&& f)
{
    using std::integral_constant;
#define HPCGX_PAIR(A, B)                                                  \
    if(wl <= (A) && wt <= (B))                                            \
    {                                                                     \
        f(integral_constant<int, A>{}, integral_constant<int, B>{});      \
        return;                                                           \
    }
    HPCGX_PAIR(0, 8)
    HPCGX_PAIR(2, 4)
    HPCGX_PAIR(10, 2)
    HPCGX_PAIR(12, 2)
    HPCGX_PAIR(16, 2)
    HPCGX_PAIR(16, 4)
    HPCGX_PAIR(22, 4)
    HPCGX_PAIR(W - 1, 8)
#undef HPCGX_PAIR
    f(integral_constant<int, W - 1>{}, integral_constant<int, W - 1>{});
}

}  // namespace hpcgx
