// ops.hip -- the solver's hot-path kernels for gfx950 (CDNA4, wave64).
//
// All operators are HBM-bound sparse/streaming work: one thread per row,
// column-major ELL (entry p of row r at p*m+r) so every slot load of a wave is
// one contiguous 256 B (cols) / 512 B (f64 values) segment, matrix streams
// loaded non-temporally, x gathered through L2/MALL.  Row tiles are laid out
// so each XCD works on one contiguous spatial slab (xcd_block), which keeps
// the neighbours a slab gathers inside that XCD's L2.
//
// Reference kernels whose results these reproduce:
//   SpMV                kernel_spmv_ell           ComputeSPMV.cpp:130-170
//   SymGS sweep         kernel_symgs_sweep        ComputeSYMGS.cpp:113-151, :369-404
//   zero-guess SymGS    kernel_pointwise_mult / kernel_forward_sweep_0 /
//                       kernel_backward_sweep_0   ComputeSYMGS.cpp:238-341, :406-454
//   residual+restrict   kernel_fused_restrict_spmv ComputeRestriction.cpp:81-124
//   prolongation        kernel_prolongation       ComputeProlongation.cpp:39-59
//   dot / waxpby / fused waxpby-dot               ComputeDotProduct.cpp:63-238,
//                                                 ComputeWAXPBY.cpp:63-240
#include "common.hpp"
#include "ops.hpp"

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
// deterministic block-partials reduction; the last block folds the partials
// in index order, applies `op` and re-arms the ticket.
__device__ void apply_op(int op, double v, CGScalars* sc, double* hist, int hist_cap, double* out)
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

template <int BS>
__device__ __forceinline__ void reduce_finish(double v, const RedArgs& ra)
{
    __shared__ double lds[BS / 64];
    __shared__ int is_last;
    const double bsum = block_sum<BS>(v, lds);
    if(threadIdx.x == 0)
    {
        ra.partial[blockIdx.x] = bsum;
        __threadfence();
        const unsigned t = atomicAdd(ra.ticket, 1u);
        is_last = (t == gridDim.x - 1) ? 1 : 0;
    }
    __syncthreads();
    if(!is_last) return;
    __threadfence();
    double s = 0.0;
    for(unsigned i = threadIdx.x; i < gridDim.x; i += BS) s += __builtin_nontemporal_load(ra.partial + i);
    __syncthreads();
    s = block_sum<BS>(s, lds);
    if(threadIdx.x == 0)
    {
        apply_op(ra.op, s, ra.sc, ra.hist, ra.hist_cap, ra.out);
        atomicExch(ra.ticket, 0u);
    }
}

// ---------------------------------------------------------------------------
// Row tiles of a whole-level operator: tile t -> (slab j = t / ncolors,
// colour c = t % ncolors), rows offsets[c] + j*BS + [0, BS) clipped to the
// colour.  Consecutive tiles therefore cover the same spatial slab in every
// colour, and each block takes a contiguous run of `tpb` tiles of its XCD's
// slice.
struct Tiles
{
    int ncolors;
    int jtiles;   // max over colours of ceil(size/BS)
    int tpb;      // tiles per block
    int off[MAX_COLORS];
    int size[MAX_COLORS];
};

// Row loads are issued in phases -- all slot indices and values of the row
// first (independent, non-temporal), then the x gathers of the valid slots,
// then the fma chain in ascending slot order -- so one row keeps 27+27+27
// requests in flight instead of serialising slot by slot.
template <typename VT, typename XT>
__device__ __forceinline__ void load_row(const int32_t* cp, const VT* vp, int64_t m, int32_t (&cc)[W], XT (&a)[W])
{
#pragma unroll
    for(int p = 0; p < W; ++p) cc[p] = ntload(cp + (int64_t)p * m);
#pragma unroll
    for(int p = 0; p < W; ++p) a[p] = cvt<XT>(ntload(vp + (int64_t)p * m));
}

template <typename XT>
__device__ __forceinline__ void gather(const XT* __restrict__ x, const int32_t (&cc)[W], XT (&xv)[W])
{
#pragma unroll
    for(int p = 0; p < W; ++p) xv[p] = (cc[p] >= 0) ? x[cc[p]] : (XT)0;
}

template <typename VT, typename XT, typename YT, bool DOT, int BS>
__global__ __launch_bounds__(BS) void k_spmv(int64_t m, const int32_t* __restrict__ col,
                                            const VT* __restrict__ val, const XT* __restrict__ x,
                                            YT* __restrict__ y, Tiles tl, RedArgs ra, const int* gate)
{
    if(gated(gate)) return;
    const int64_t L = xcd_block(blockIdx.x, gridDim.x);
    const int64_t ntiles = (int64_t)tl.ncolors * tl.jtiles;
    const int64_t t0 = L * tl.tpb;
    const int64_t t1 = (t0 + tl.tpb < ntiles) ? t0 + tl.tpb : ntiles;
    double acc = 0.0;
    for(int64_t t = t0; t < t1; ++t)
    {
        const int c = (int)(t % tl.ncolors);
        const int64_t i = (t / tl.ncolors) * BS + threadIdx.x;
        if(i >= tl.size[c]) continue;
        const int64_t row = tl.off[c] + i;
        int32_t cc[W];
        XT a[W], xv[W];
        load_row(col + row, val + row, m, cc, a);
        gather(x, cc, xv);
        XT s = 0;
#pragma unroll
        for(int p = 0; p < W; ++p)
            if(cc[p] >= 0) s = fmaT(a[p], xv[p], s);
        y[row] = (YT)s;
        if(DOT) acc = fmaT((double)x[row], (double)s, acc);
    }
    if(DOT) reduce_finish<BS>(acc, ra);
}

// Zero-initial-guess forward sweep of colour c: only entries left of the
// diagonal (earlier colours) are non-zero in x.
template <typename VT, typename XT, typename RT, int BS>
__global__ __launch_bounds__(BS) void k_fwd_zero(int64_t m, int64_t off, int64_t size, const int32_t* __restrict__ col,
                                                 const VT* __restrict__ val, const int32_t* __restrict__ dgi,
                                                 const RT* __restrict__ r, XT* __restrict__ x, const int* gate)
{
    if(gated(gate)) return;
    const int64_t i = xcd_block(blockIdx.x, gridDim.x) * BS + threadIdx.x;
    if(i >= size) return;
    const int64_t row = off + i;
    const int dg = dgi[row];
    const int32_t* cp = col + row;
    const VT* vp = val + row;
    int32_t cc[W];
    XT a[W], xv[W];
#pragma unroll
    for(int p = 0; p < W; ++p) cc[p] = (p < dg) ? ntload(cp + (int64_t)p * m) : -1;
#pragma unroll
    for(int p = 0; p < W; ++p) a[p] = (p < dg) ? cvt<XT>(ntload(vp + (int64_t)p * m)) : (XT)0;
    const XT d = cvt<XT>(vp[(int64_t)dg * m]);
    XT s = cvt<XT>(r[row]);
    gather(x, cc, xv);
#pragma unroll
    for(int p = 0; p < W; ++p)
        if(p < dg) s = fmaT(-a[p], xv[p], s);
    x[row] = s / d;
}

// Zero-initial-guess backward sweep of colour c: x = (x*d - U x) / d.
template <typename VT, typename XT, int BS>
__global__ __launch_bounds__(BS) void k_bwd_zero(int64_t m, int64_t off, int64_t size, const int32_t* __restrict__ col,
                                                 const VT* __restrict__ val, const int32_t* __restrict__ dgi,
                                                 XT* __restrict__ x, const int* gate)
{
    if(gated(gate)) return;
    const int64_t i = xcd_block(blockIdx.x, gridDim.x) * BS + threadIdx.x;
    if(i >= size) return;
    const int64_t row = off + i;
    const int dg = dgi[row];
    const int32_t* cp = col + row;
    const VT* vp = val + row;
    int32_t cc[W];
    XT a[W], xv[W];
#pragma unroll
    for(int p = 0; p < W; ++p) cc[p] = (p > dg) ? ntload(cp + (int64_t)p * m) : -1;
#pragma unroll
    for(int p = 0; p < W; ++p) a[p] = (p > dg) ? cvt<XT>(ntload(vp + (int64_t)p * m)) : (XT)0;
    const XT d = cvt<XT>(vp[(int64_t)dg * m]);
    XT s = x[row] * d;
    gather(x, cc, xv);
#pragma unroll
    for(int p = 0; p < W; ++p)
        if(cc[p] >= 0) s = fmaT(-a[p], xv[p], s);
    x[row] = s / d;
}

// Full forward sweep of colour c over the whole row, in ascending column
// order; keeps t = r - L x (the sum up to the diagonal) for the backward
// sweep, whose lower neighbours do not change in between.
template <typename VT, typename XT, typename RT, bool STORE_T, int BS>
__global__ __launch_bounds__(BS) void k_fwd_full(int64_t m, int64_t off, int64_t size, const int32_t* __restrict__ col,
                                                 const VT* __restrict__ val, const RT* __restrict__ r,
                                                 XT* __restrict__ x, XT* __restrict__ t, const int* gate)
{
    if(gated(gate)) return;
    const int64_t i = xcd_block(blockIdx.x, gridDim.x) * BS + threadIdx.x;
    if(i >= size) return;
    const int64_t row = off + i;
    int32_t cc[W];
    XT a[W], xv[W];
    load_row(col + row, val + row, m, cc, a);
#pragma unroll
    for(int p = 0; p < W; ++p) xv[p] = (cc[p] >= 0 && (int64_t)cc[p] != row) ? x[cc[p]] : (XT)0;
    XT s = cvt<XT>(r[row]);
    XT tt = 0, d = 1;
#pragma unroll
    for(int p = 0; p < W; ++p)
    {
        if(cc[p] >= 0)
        {
            if((int64_t)cc[p] == row)
            {
                tt = s;
                d = a[p];
            }
            else
            {
                s = fmaT(-a[p], xv[p], s);
            }
        }
    }
    x[row] = s / d;
    if(STORE_T) t[row] = tt;
}

// Full backward sweep of colour c using t: x = (t - U x) / d.
template <typename VT, typename XT, int BS>
__global__ __launch_bounds__(BS) void k_bwd_full(int64_t m, int64_t off, int64_t size, const int32_t* __restrict__ col,
                                                 const VT* __restrict__ val, const int32_t* __restrict__ dgi,
                                                 const XT* __restrict__ t, XT* __restrict__ x, const int* gate)
{
    if(gated(gate)) return;
    const int64_t i = xcd_block(blockIdx.x, gridDim.x) * BS + threadIdx.x;
    if(i >= size) return;
    const int64_t row = off + i;
    const int dg = dgi[row];
    const int32_t* cp = col + row;
    const VT* vp = val + row;
    int32_t cc[W];
    XT a[W], xv[W];
#pragma unroll
    for(int p = 0; p < W; ++p) cc[p] = (p > dg) ? ntload(cp + (int64_t)p * m) : -1;
#pragma unroll
    for(int p = 0; p < W; ++p) a[p] = (p > dg) ? cvt<XT>(ntload(vp + (int64_t)p * m)) : (XT)0;
    const XT d = cvt<XT>(vp[(int64_t)dg * m]);
    XT s = t[row];
    gather(x, cc, xv);
#pragma unroll
    for(int p = 0; p < W; ++p)
        if(cc[p] >= 0) s = fmaT(-a[p], xv[p], s);
    x[row] = s / d;
}

// Fused fine residual + injection for coarse natural row k:
// rc[f2c_coarse[k]] = r[f] - sum_p a[f,p] x[col], f = f2c_fine[k].
template <typename VT, typename XT, typename RT, typename CT, int BS>
__global__ __launch_bounds__(BS) void k_restrict(int64_t m, int64_t mc, const int32_t* __restrict__ col,
                                                 const VT* __restrict__ val, const int32_t* __restrict__ f2c_fine,
                                                 const int32_t* __restrict__ f2c_coarse, const RT* __restrict__ r,
                                                 const XT* __restrict__ x, CT* __restrict__ rc, const int* gate)
{
    if(gated(gate)) return;
    const int64_t k = xcd_block(blockIdx.x, gridDim.x) * BS + threadIdx.x;
    if(k >= mc) return;
    const int64_t row = f2c_fine[k];
    int32_t cc[W];
    XT a[W], xv[W];
    load_row(col + row, val + row, m, cc, a);
    gather(x, cc, xv);
    XT s = cvt<XT>(r[row]);
#pragma unroll
    for(int p = 0; p < W; ++p)
        if(cc[p] >= 0) s = fmaT(-a[p], xv[p], s);
    rc[f2c_coarse[k]] = (CT)s;
}

// Prolongation: xf[f2c_fine[k]] += xc[f2c_coarse[k]].
template <typename XT, typename CT, int BS>
__global__ __launch_bounds__(BS) void k_prolong(int64_t mc, const int32_t* __restrict__ f2c_fine,
                                                const int32_t* __restrict__ f2c_coarse, const CT* __restrict__ xc,
                                                XT* __restrict__ xf, const int* gate)
{
    if(gated(gate)) return;
    const int64_t k = xcd_block(blockIdx.x, gridDim.x) * BS + threadIdx.x;
    if(k >= mc) return;
    const int64_t f = f2c_fine[k];
    xf[f] = xf[f] + (XT)xc[f2c_coarse[k]];
}

// ---------------------------------------------------------------------------
// CG vector kernels (f64 vectors; z may be f32 in the mixed mode).  Chunked:
// block L (XCD-contiguous) handles elements [L*chunk, (L+1)*chunk).
__device__ __forceinline__ void chunk_range(int64_t n, int64_t& b, int64_t& e)
{
    const int64_t L = xcd_block(blockIdx.x, gridDim.x);
    const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
    b = L * chunk;
    e = (b + chunk < n) ? b + chunk : n;
}

template <typename XT, typename YT, int BS>
__global__ __launch_bounds__(BS) void k_dot(int64_t n, const XT* __restrict__ x, const YT* __restrict__ y,
                                            RedArgs ra, const int* gate)
{
    if(gated(gate)) return;
    int64_t b, e;
    chunk_range(n, b, e);
    double acc = 0.0;
    for(int64_t i = b + threadIdx.x; i < e; i += BS) acc = fmaT((double)y[i], (double)x[i], acc);
    reduce_finish<BS>(acc, ra);
}

// w = alpha*x + beta*y with the reference's special cases (ComputeWAXPBY.cpp:63-91)
template <int BS>
__global__ __launch_bounds__(BS) void k_waxpby(int64_t n, double alpha, const double* __restrict__ x, double beta,
                                               const double* __restrict__ y, double* __restrict__ w, const int* gate)
{
    if(gated(gate)) return;
    int64_t b, e;
    chunk_range(n, b, e);
    for(int64_t i = b + threadIdx.x; i < e; i += BS)
    {
        if(alpha == 1.0)
            w[i] = fmaT(beta, y[i], x[i]);
        else if(beta == 1.0)
            w[i] = fmaT(alpha, x[i], y[i]);
        else
            w[i] = fmaT(alpha, x[i], beta * y[i]);
    }
}

// y += alpha*x; result = y'y (ComputeWAXPBY.cpp:137-240)
template <int BS>
__global__ __launch_bounds__(BS) void k_axpy_dot(int64_t n, double alpha, const double* __restrict__ x,
                                                 double* __restrict__ y, RedArgs ra, const int* gate)
{
    if(gated(gate)) return;
    int64_t b, e;
    chunk_range(n, b, e);
    double acc = 0.0;
    for(int64_t i = b + threadIdx.x; i < e; i += BS)
    {
        const double v = fmaT(alpha, x[i], y[i]);
        y[i] = v;
        acc = fmaT(v, v, acc);
    }
    reduce_finish<BS>(acc, ra);
}

// p = z (first iteration) or p = z + beta*p, with beta read on the device
template <typename ZT, int BS>
__global__ __launch_bounds__(BS) void k_cg_p(int64_t n, const ZT* __restrict__ z, double* __restrict__ p,
                                             const CGScalars* __restrict__ sc, const int* gate)
{
    if(gated(gate)) return;
    int64_t b, e;
    chunk_range(n, b, e);
    const bool first = sc->iter == 0;
    const double beta = sc->beta;
    for(int64_t i = b + threadIdx.x; i < e; i += BS)
        p[i] = first ? (double)z[i] : fmaT(beta, p[i], (double)z[i]);
}

// x += alpha*p; r -= alpha*Ap; ||r||^2 -> normr, with alpha read on the device
template <int BS>
__global__ __launch_bounds__(BS) void k_cg_xr(int64_t n, const double* __restrict__ p, const double* __restrict__ Ap,
                                              double* __restrict__ x, double* __restrict__ r, RedArgs ra,
                                              const int* gate)
{
    if(gated(gate)) return;
    int64_t b, e;
    chunk_range(n, b, e);
    const double alpha = ra.sc->alpha;
    double acc = 0.0;
    for(int64_t i = b + threadIdx.x; i < e; i += BS)
    {
        const double rv = fmaT(-alpha, Ap[i], r[i]);
        r[i] = rv;
        x[i] = fmaT(alpha, p[i], x[i]);
        acc = fmaT(rv, rv, acc);
    }
    reduce_finish<BS>(acc, ra);
}

template <typename T, int BS>
__global__ __launch_bounds__(BS) void k_fill(int64_t n, T* __restrict__ p, T v, const int* gate)
{
    if(gated(gate)) return;
    const int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
    if(i < n) p[i] = v;
}

template <typename T, typename U, int BS>
__global__ __launch_bounds__(BS) void k_copy(int64_t n, const T* __restrict__ src, U* __restrict__ dst, const int* gate)
{
    if(gated(gate)) return;
    const int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
    if(i < n) dst[i] = (U)src[i];
}

// Diagonal access for TestCG (SparseMatrix.cpp:66-152)
template <typename VT>
__global__ void k_get_diag(int64_t m, const int32_t* __restrict__ dgi, const VT* __restrict__ val, double* __restrict__ d)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if(r < m) d[r] = cvt<double>(val[(int64_t)dgi[r] * m + r]);
}

__global__ void k_set_diag(int64_t m, const int32_t* __restrict__ dgi, double* __restrict__ val,
                           double* __restrict__ inv_diag, const double* __restrict__ d)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if(r < m)
    {
        val[(int64_t)dgi[r] * m + r] = d[r];
        if(inv_diag) inv_diag[r] = 1.0 / d[r];
    }
}

// reference TestCG scaling (TestCG.cpp:67-95): position i keeps the factor of
// global row i because the reference never permutes d_localToGlobalMap.
__global__ void k_testcg_scale(int64_t m, double* __restrict__ d, double* __restrict__ b)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= m) return;
    const double scale = (i < 9) ? (i + 2) * 1.0e6 : 1.0e6;
    d[i] *= scale;
    b[i] *= scale;
}

// ---------------------------------------------------------------------------
// host launchers
static inline unsigned nblk(int64_t n, int bs) { return (unsigned)ceil_div(n, bs); }

#define HPCGX_KCHECK() HPCGX_HIP(hipGetLastError())

static Tiles make_tiles(const Level& L, int bs, int max_blocks, int64_t& nblocks)
{
    Tiles t{};
    t.ncolors = L.ncolors;
    int mx = 0;
    for(int c = 0; c < L.ncolors; ++c)
    {
        t.off[c] = L.offsets[c];
        t.size[c] = L.sizes[c];
        mx = mx > L.sizes[c] ? mx : L.sizes[c];
    }
    t.jtiles = (int)ceil_div(mx, bs);
    const int64_t ntiles = (int64_t)t.ncolors * t.jtiles;
    t.tpb = (max_blocks > 0) ? (int)ceil_div(ntiles, max_blocks) : 1;
    if(t.tpb < 1) t.tpb = 1;
    nblocks = ceil_div(ntiles, t.tpb);
    return t;
}

// element-type views of a level's arrays
template <typename VT>
static const VT* vals_of(const Level& L);
template <>
const double* vals_of<double>(const Level& L) { return L.ell_val; }
template <>
const __half* vals_of<__half>(const Level& L) { return L.ell_h; }

template <typename VT, typename XT, typename YT>
static int spmv_t(const Level& L, const void* x, void* y, const RedArgs* ra, hipStream_t s, const int* gate)
{
    constexpr int BS = 256;
    int64_t nb;
    Tiles tl = make_tiles(L, BS, ra ? RED_BLOCKS : 0, nb);
    if(ra)
        hipLaunchKernelGGL((k_spmv<VT, XT, YT, true, BS>), dim3((unsigned)nb), dim3(BS), 0, s, L.m, L.ell_col,
                           vals_of<VT>(L), (const XT*)x, (YT*)y, tl, *ra, gate);
    else
        hipLaunchKernelGGL((k_spmv<VT, XT, YT, false, BS>), dim3((unsigned)nb), dim3(BS), 0, s, L.m, L.ell_col,
                           vals_of<VT>(L), (const XT*)x, (YT*)y, tl, RedArgs{}, gate);
    HPCGX_KCHECK();
    return 0;
}

int launch_spmv(const Level& L, Prec prec, const void* x, void* y, hipStream_t s, const int* gate)
{
    if(prec == PREC_F64) return spmv_t<double, double, double>(L, x, y, nullptr, s, gate);
    return spmv_t<__half, float, float>(L, x, y, nullptr, s, gate);
}

int launch_spmv_f64x(const Level& L, Prec prec, const double* x, double* y, const RedArgs* ra, hipStream_t s,
                     const int* gate)
{
    // f64 vectors over either value store (f16 values are exact: 26 / -1)
    if(prec == PREC_F64) return spmv_t<double, double, double>(L, x, y, ra, s, gate);
    return spmv_t<__half, double, double>(L, x, y, ra, s, gate);
}

template <typename VT, typename XT, typename RT>
static int symgs_zero_t(const Level& L, const void* rv, void* xv, hipStream_t s, const int* gate)
{
    constexpr int BS = 256;
    const RT* r = (const RT*)rv;
    XT* x = (XT*)xv;
    const VT* val = vals_of<VT>(L);
    for(int c = 0; c < L.ncolors; ++c)
    {
        if(L.sizes[c] == 0) continue;
        hipLaunchKernelGGL((k_fwd_zero<VT, XT, RT, BS>), dim3(nblk(L.sizes[c], BS)), dim3(BS), 0, s, L.m,
                           (int64_t)L.offsets[c], (int64_t)L.sizes[c], L.ell_col, val, L.diag_idx, r, x, gate);
        HPCGX_KCHECK();
    }
    for(int c = L.ncolors - 2; c >= 0; --c)
    {
        if(L.sizes[c] == 0) continue;
        hipLaunchKernelGGL((k_bwd_zero<VT, XT, BS>), dim3(nblk(L.sizes[c], BS)), dim3(BS), 0, s, L.m,
                           (int64_t)L.offsets[c], (int64_t)L.sizes[c], L.ell_col, val, L.diag_idx, x, gate);
        HPCGX_KCHECK();
    }
    return 0;
}

template <typename VT, typename XT, typename RT>
static int symgs_full_t(const Level& L, const void* rv, void* xv, void* tv, hipStream_t s, const int* gate)
{
    constexpr int BS = 256;
    const RT* r = (const RT*)rv;
    XT* x = (XT*)xv;
    XT* t = (XT*)tv;
    const VT* val = vals_of<VT>(L);
    const int last = L.ncolors - 1;
    for(int c = 0; c < L.ncolors; ++c)
    {
        if(L.sizes[c] == 0) continue;
        if(c < last)
            hipLaunchKernelGGL((k_fwd_full<VT, XT, RT, true, BS>), dim3(nblk(L.sizes[c], BS)), dim3(BS), 0, s, L.m,
                               (int64_t)L.offsets[c], (int64_t)L.sizes[c], L.ell_col, val, r, x, t, gate);
        else
            hipLaunchKernelGGL((k_fwd_full<VT, XT, RT, false, BS>), dim3(nblk(L.sizes[c], BS)), dim3(BS), 0, s,
                               L.m, (int64_t)L.offsets[c], (int64_t)L.sizes[c], L.ell_col, val, r, x, t, gate);
        HPCGX_KCHECK();
    }
    for(int c = last - 1; c >= 0; --c)
    {
        if(L.sizes[c] == 0) continue;
        hipLaunchKernelGGL((k_bwd_full<VT, XT, BS>), dim3(nblk(L.sizes[c], BS)), dim3(BS), 0, s, L.m,
                           (int64_t)L.offsets[c], (int64_t)L.sizes[c], L.ell_col, val, L.diag_idx, (const XT*)t, x,
                           gate);
        HPCGX_KCHECK();
    }
    return 0;
}

int launch_symgs_zero(const Level& L, Prec prec, const void* r, bool r_f64, void* x, hipStream_t s, const int* gate)
{
    if(prec == PREC_F64) return symgs_zero_t<double, double, double>(L, r, x, s, gate);
    if(r_f64) return symgs_zero_t<__half, float, double>(L, r, x, s, gate);
    return symgs_zero_t<__half, float, float>(L, r, x, s, gate);
}

int launch_symgs(const Level& L, Prec prec, const void* r, bool r_f64, void* x, void* t, hipStream_t s,
                 const int* gate)
{
    if(prec == PREC_F64) return symgs_full_t<double, double, double>(L, r, x, t, s, gate);
    if(r_f64) return symgs_full_t<__half, float, double>(L, r, x, t, s, gate);
    return symgs_full_t<__half, float, float>(L, r, x, t, s, gate);
}

template <typename VT, typename XT, typename RT>
static int restrict_t(const Level& F, const void* r, const void* x, void* rc, hipStream_t s, const int* gate)
{
    constexpr int BS = 256;
    hipLaunchKernelGGL((k_restrict<VT, XT, RT, XT, BS>), dim3(nblk(F.mc, BS)), dim3(BS), 0, s, F.m, F.mc,
                       F.ell_col, vals_of<VT>(F), F.f2c_fine, F.f2c_coarse, (const RT*)r, (const XT*)x, (XT*)rc,
                       gate);
    HPCGX_KCHECK();
    return 0;
}

int launch_restrict(const Level& F, Prec prec, const void* r, bool r_f64, const void* x, void* rc, hipStream_t s,
                    const int* gate)
{
    if(prec == PREC_F64) return restrict_t<double, double, double>(F, r, x, rc, s, gate);
    if(r_f64) return restrict_t<__half, float, double>(F, r, x, rc, s, gate);
    return restrict_t<__half, float, float>(F, r, x, rc, s, gate);
}

int launch_prolong(const Level& F, Prec prec, const void* xc, void* x, hipStream_t s, const int* gate)
{
    constexpr int BS = 256;
    if(prec == PREC_F64)
        hipLaunchKernelGGL((k_prolong<double, double, BS>), dim3(nblk(F.mc, BS)), dim3(BS), 0, s, F.mc, F.f2c_fine,
                           F.f2c_coarse, (const double*)xc, (double*)x, gate);
    else
        hipLaunchKernelGGL((k_prolong<float, float, BS>), dim3(nblk(F.mc, BS)), dim3(BS), 0, s, F.mc, F.f2c_fine,
                           F.f2c_coarse, (const float*)xc, (float*)x, gate);
    HPCGX_KCHECK();
    return 0;
}

static unsigned red_grid(int64_t n)
{
    int64_t g = ceil_div(n, (int64_t)RED_THREADS * 4);
    if(g > RED_BLOCKS) g = RED_BLOCKS;
    if(g < 1) g = 1;
    return (unsigned)g;
}

int launch_dot(int64_t n, const void* x, bool x_f64, const void* y, bool y_f64, const RedArgs& ra, hipStream_t s,
               const int* gate)
{
    constexpr int BS = RED_THREADS;
    const unsigned g = red_grid(n);
    if(x_f64 && y_f64)
        hipLaunchKernelGGL((k_dot<double, double, BS>), dim3(g), dim3(BS), 0, s, n, (const double*)x,
                           (const double*)y, ra, gate);
    else if(x_f64 && !y_f64)
        hipLaunchKernelGGL((k_dot<double, float, BS>), dim3(g), dim3(BS), 0, s, n, (const double*)x,
                           (const float*)y, ra, gate);
    else if(!x_f64 && y_f64)
        hipLaunchKernelGGL((k_dot<float, double, BS>), dim3(g), dim3(BS), 0, s, n, (const float*)x,
                           (const double*)y, ra, gate);
    else
        hipLaunchKernelGGL((k_dot<float, float, BS>), dim3(g), dim3(BS), 0, s, n, (const float*)x, (const float*)y,
                           ra, gate);
    HPCGX_KCHECK();
    return 0;
}

int launch_waxpby(int64_t n, double alpha, const double* x, double beta, const double* y, double* w, hipStream_t s,
                  const int* gate)
{
    constexpr int BS = RED_THREADS;
    hipLaunchKernelGGL((k_waxpby<BS>), dim3(red_grid(n)), dim3(BS), 0, s, n, alpha, x, beta, y, w, gate);
    HPCGX_KCHECK();
    return 0;
}

int launch_axpy_dot(int64_t n, double alpha, const double* x, double* y, const RedArgs& ra, hipStream_t s,
                    const int* gate)
{
    constexpr int BS = RED_THREADS;
    hipLaunchKernelGGL((k_axpy_dot<BS>), dim3(red_grid(n)), dim3(BS), 0, s, n, alpha, x, y, ra, gate);
    HPCGX_KCHECK();
    return 0;
}

int launch_cg_p(int64_t n, const void* z, bool z_f64, double* p, const CGScalars* sc, hipStream_t s, const int* gate)
{
    constexpr int BS = RED_THREADS;
    if(z_f64)
        hipLaunchKernelGGL((k_cg_p<double, BS>), dim3(red_grid(n)), dim3(BS), 0, s, n, (const double*)z, p, sc, gate);
    else
        hipLaunchKernelGGL((k_cg_p<float, BS>), dim3(red_grid(n)), dim3(BS), 0, s, n, (const float*)z, p, sc, gate);
    HPCGX_KCHECK();
    return 0;
}

int launch_cg_xr(int64_t n, const double* p, const double* Ap, double* x, double* r, const RedArgs& ra,
                 hipStream_t s, const int* gate)
{
    constexpr int BS = RED_THREADS;
    hipLaunchKernelGGL((k_cg_xr<BS>), dim3(red_grid(n)), dim3(BS), 0, s, n, p, Ap, x, r, ra, gate);
    HPCGX_KCHECK();
    return 0;
}

int launch_fill(int64_t n, void* p, bool f64, double v, hipStream_t s, const int* gate)
{
    if(n <= 0) return 0;
    if(f64)
        hipLaunchKernelGGL((k_fill<double, 256>), dim3(nblk(n, 256)), dim3(256), 0, s, n, (double*)p, v, gate);
    else
        hipLaunchKernelGGL((k_fill<float, 256>), dim3(nblk(n, 256)), dim3(256), 0, s, n, (float*)p, (float)v, gate);
    HPCGX_KCHECK();
    return 0;
}

int launch_copy(int64_t n, const void* src, bool src_f64, void* dst, bool dst_f64, hipStream_t s, const int* gate)
{
    if(n <= 0) return 0;
    const dim3 g(nblk(n, 256)), b(256);
    if(src_f64 && dst_f64)
        hipLaunchKernelGGL((k_copy<double, double, 256>), g, b, 0, s, n, (const double*)src, (double*)dst, gate);
    else if(src_f64)
        hipLaunchKernelGGL((k_copy<double, float, 256>), g, b, 0, s, n, (const double*)src, (float*)dst, gate);
    else if(dst_f64)
        hipLaunchKernelGGL((k_copy<float, double, 256>), g, b, 0, s, n, (const float*)src, (double*)dst, gate);
    else
        hipLaunchKernelGGL((k_copy<float, float, 256>), g, b, 0, s, n, (const float*)src, (float*)dst, gate);
    HPCGX_KCHECK();
    return 0;
}

int launch_get_diag(const Level& L, double* d, hipStream_t s)
{
    if(L.ell_val)
        hipLaunchKernelGGL(k_get_diag<double>, dim3(nblk(L.m, 256)), dim3(256), 0, s, L.m, L.diag_idx, L.ell_val, d);
    else
        hipLaunchKernelGGL(k_get_diag<__half>, dim3(nblk(L.m, 256)), dim3(256), 0, s, L.m, L.diag_idx, L.ell_h, d);
    HPCGX_KCHECK();
    return 0;
}

int launch_set_diag(Level& L, const double* d, hipStream_t s)
{
    if(!L.ell_val)
    {
        set_error("diagonal replacement needs an f64 level");
        return -2;
    }
    hipLaunchKernelGGL(k_set_diag, dim3(nblk(L.m, 256)), dim3(256), 0, s, L.m, L.diag_idx, L.ell_val, L.inv_diag, d);
    HPCGX_KCHECK();
    return 0;
}

int launch_testcg_scale(int64_t m, double* d, double* b, hipStream_t s)
{
    hipLaunchKernelGGL(k_testcg_scale, dim3(nblk(m, 256)), dim3(256), 0, s, m, d, b);
    HPCGX_KCHECK();
    return 0;
}

}  // namespace hpcgx
