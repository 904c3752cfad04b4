// ops.hip -- the solver's hot-path kernels for gfx950 (CDNA4, wave64).
//
// All operators are HBM-bound sparse/streaming work: one thread per row,
// sliced ELL (entry p of row r at ell_at(p, r), common.hpp) so every slot load
// of a wave is one contiguous 256 B (cols) / 512 B (f64 values) segment and
// the slots of a slice follow each other, matrix streams loaded non-temporally, x gathered through L2/MALL.  Row tiles are laid out
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
#include "kernels.hpp"
#include "ops.hpp"

#include <algorithm>
#include <type_traits>

namespace hpcgx
{

template <typename VT, typename XT, typename YT, bool DOT, int R, int BS>
__global__ __launch_bounds__(BS) void k_spmv(const int32_t* __restrict__ col,
                                            const VT* __restrict__ val, const XT* __restrict__ x,
                                            YT* __restrict__ y, Tiles tl, RedArgs ra, const int* gate)
{
    if(gated(gate)) return;
    const int64_t ntiles = (int64_t)tl.ncolors * tl.jtiles;
    const int64_t L = xcd_block(blockIdx.x, gridDim.x);
    const int64_t g0 = (L / tl.gx) * tl.gx;   // first block of L's group
    const int64_t members = (g0 + tl.gx < (int64_t)gridDim.x) ? tl.gx : (int64_t)gridDim.x - g0;
    const int64_t e = (g0 + members) * tl.tpb;
    const int64_t t0 = g0 * tl.tpb + (L - g0);
    const int64_t t1 = e < ntiles ? e : ntiles;
    double acc = 0.0;
    for(int64_t t = t0; t < t1; t += members)
    {
        const int c = (int)(t % tl.ncolors);
        const int64_t i = (t / tl.ncolors) * (BS * R) + (int64_t)threadIdx.x * R;
        if(i >= tl.size[c]) continue;
        const int64_t row = tl.off[c] + i;
        XT s[R];
        row_spmv<VT, XT, R>(row, col, val, x, s);
        st_cvt<R>(y + row, s);
        if(DOT)
        {
            XT xr[R];
            ld<R>(x + row, xr);
#pragma unroll
            for(int k = 0; k < R; ++k) acc = fmaT((double)xr[k], (double)s[k], acc);
        }
    }
    if(DOT) block_partial<BS>(acc, ra.partial, blockIdx.x);
}

// Per-colour sweeps: block b covers rows off + T*BS*R + [0, BS*R), where T is
// L = xcd_block(b), mapped through the colour's column order (ZMap) on large
// planes.  Consecutive colour launches alternate direction (size < 0 marks a
// reversed launch): each XCD then starts a launch on the slab it finished the
// previous one on, whose neighbours' x is still in its caches.
#define SWEEP_ROW()                                                                   \
    if(gated(gate)) return;                                                           \
    int64_t Lb_ = xcd_block(blockIdx.x, gridDim.x);                                   \
    if(size < 0)                                                                      \
    {                                                                                 \
        size = -size;                                                                 \
        Lb_ = (int64_t)gridDim.x - 1 - Lb_;                                           \
    }                                                                                 \
    if(zm.k > 1)                                                                      \
    {                                                                                 \
        const unsigned per_ = (unsigned)zm.planes * (unsigned)zm.tpc;                 \
        const unsigned lb_ = (unsigned)Lb_, col_ = lb_ / per_, rem_ = lb_ - col_ * per_; \
        const unsigned pl_ = rem_ / (unsigned)zm.tpc;                                 \
        Lb_ = ((int64_t)pl_ * zm.k + col_) * zm.tpc + (rem_ - pl_ * (unsigned)zm.tpc); \
    }                                                                                 \
    const int64_t i = Lb_ * (BS * R) + (int64_t)threadIdx.x * R;                      \
    if(i >= size) return;                                                             \
    const int64_t row = off + i;

template <int NS, typename VT, typename XT, typename RT, int R, int BS>
__global__ __launch_bounds__(BS) void k_fwd_zero(int64_t off, int64_t size, ZMap zm, const int32_t* __restrict__ col, const VT* __restrict__ val,
                                                 const VT* __restrict__ dg, const int32_t* __restrict__ desc,
                                                 const RT* __restrict__ r, XT* __restrict__ x, const int* gate)
{
    SWEEP_ROW();
    row_fwd_zero<NS, VT, XT, RT, R>(row, col, val, dg, desc, r, x);
}

template <int NS, typename VT, typename XT, bool STORE_U, int R, int BS>
__global__ __launch_bounds__(BS) void k_bwd_zero(int64_t off, int64_t size, ZMap zm, const int32_t* __restrict__ col, const VT* __restrict__ val,
                                                 const VT* __restrict__ dg, const int32_t* __restrict__ desc,
                                                 XT* __restrict__ x, XT* __restrict__ u, const int* gate)
{
    SWEEP_ROW();
    row_bwd_zero<NS, VT, XT, STORE_U, R>(row, col, val, dg, desc, x, u);
}

template <typename VT, typename XT, typename RT, bool STORE_T, int R, int BS>
__global__ __launch_bounds__(BS) void k_fwd_full(int64_t off, int64_t size, ZMap zm, const int32_t* __restrict__ col,
                                                 const VT* __restrict__ val, const RT* __restrict__ r,
                                                 XT* __restrict__ x, XT* __restrict__ t, const int* gate)
{
    SWEEP_ROW();
    row_fwd_full<VT, XT, RT, STORE_T, R>(row, col, val, r, x, t);
}

template <int NA, int NB, typename VT, typename XT, typename RT, int R, int BS>
__global__ __launch_bounds__(BS) void k_fwd_full_u(int64_t off, int64_t size, ZMap zm,
                                                   const int32_t* __restrict__ col, const VT* __restrict__ val,
                                                   const VT* __restrict__ dg, const int32_t* __restrict__ desc,
                                                   const RT* __restrict__ r, XT* __restrict__ x, XT* __restrict__ t,
                                                   const int* gate)
{
    SWEEP_ROW();
    row_fwd_full_u<NA, NB, VT, XT, RT, R>(row, col, val, dg, desc, r, x, t);
}

template <int NS, typename VT, typename XT, int R, int BS>
__global__ __launch_bounds__(BS) void k_bwd_full(int64_t off, int64_t size, ZMap zm, const int32_t* __restrict__ col, const VT* __restrict__ val,
                                                 const VT* __restrict__ dg, const int32_t* __restrict__ desc,
                                                 const XT* __restrict__ t, XT* __restrict__ x, const int* gate)
{
    SWEEP_ROW();
    row_bwd_full<NS, VT, XT, R>(row, col, val, dg, desc, t, x);
}
#undef SWEEP_ROW

template <typename VT, typename XT, typename RT, typename CT, int BS>
__global__ __launch_bounds__(BS) void k_restrict(int64_t mc, const int32_t* __restrict__ col,
                                                 const VT* __restrict__ val, const int32_t* __restrict__ f2c_fine,
                                                 const int32_t* __restrict__ f2c_coarse, const RT* __restrict__ r,
                                                 const XT* __restrict__ x, CT* __restrict__ rc, const int* gate)
{
    if(gated(gate)) return;
    const int64_t k = xcd_block(blockIdx.x, gridDim.x) * BS + threadIdx.x;
    if(k >= mc) return;
    row_restrict<VT, XT, RT, CT>(k, col, val, f2c_fine, f2c_coarse, r, x, rc);
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

// Reference-mode pieces (ComputeSPMV.cpp:87-128 on MGData::Axf and
// ComputeRestriction.cpp:59-79): A x at the injected rows only, and
// rc = rf - Axf at those rows.
__global__ __launch_bounds__(256) void k_spmv_rows(int64_t mc, const int32_t* __restrict__ col,
                                                   const double* __restrict__ val, const int32_t* __restrict__ rows,
                                                   const double* __restrict__ x, double* __restrict__ y)
{
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if(k >= mc) return;
    const int64_t row = rows[k];
    double s = 0.0;
    for(int p = 0; p < W; ++p)
    {
        const int32_t c = col[ell_at(p, row)];
        if(c >= 0) s = fmaT(val[ell_at(p, row)], x[c], s);
    }
    y[row] = s;
}

__global__ __launch_bounds__(256) void k_restrict_diff(int64_t mc, const int32_t* __restrict__ f2c_fine,
                                                       const int32_t* __restrict__ f2c_coarse,
                                                       const double* __restrict__ rf, const double* __restrict__ Axf,
                                                       double* __restrict__ rc)
{
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if(k >= mc) return;
    const int64_t f = f2c_fine[k];
    rc[f2c_coarse[k]] = rf[f] - Axf[f];
}

// ---------------------------------------------------------------------------
// CG vector kernels (f64 vectors; z may be f32 in the mixed mode).  Chunked:
// block L (XCD-contiguous) handles elements [L*chunk, (L+1)*chunk).
// second phase of every reduction: one block folds the producer's n partials
template <int BS>
__global__ __launch_bounds__(BS) void k_fold(int64_t n, RedArgs ra, const int* gate)
{
    if(gated(gate)) return;
    fold_partials<BS>(ra.partial, n, ra);
}

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
    block_partial<BS>(acc, ra.partial, blockIdx.x);
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
    block_partial<BS>(acc, ra.partial, blockIdx.x);
}

// p = z (first iteration) or p = z + beta*p, with beta read on the device
template <typename ZT, int BS>
__global__ __launch_bounds__(BS) void k_cg_p(int64_t n, const ZT* __restrict__ z, double* __restrict__ p,
                                             double* __restrict__ x, const CGScalars* __restrict__ sc,
                                             const int* gate)
{
    // p = z + beta*p (CG.cpp:158-165), and first the previous iteration's
    // deferred x += alpha*p, which reads the p this kernel overwrites anyway
    if(gated(gate)) return;
    int64_t b, e;
    chunk_range(n, b, e);
    const bool first = sc->iter == 0;
    const bool xpend = sc->xpend != 0;
    const double beta = sc->beta;
    const double alpha = sc->alpha;
    for(int64_t i = b + threadIdx.x; i < e; i += BS)
    {
        if(first)
            p[i] = (double)z[i];
        else
        {
            const double pv = p[i];
            if(xpend) x[i] = fmaT(alpha, pv, x[i]);
            p[i] = fmaT(beta, pv, (double)z[i]);
        }
    }
}

// r -= alpha*Ap; ||r||^2 -> normr, with alpha read on the device.  The
// matching x += alpha*p is deferred to the next p update (k_cg_p) or to
// k_cg_xflush, saving one pass over p.
template <int BS>
__global__ __launch_bounds__(BS) void k_cg_xr(int64_t n, const double* __restrict__ Ap, double* __restrict__ r,
                                              RedArgs ra, const int* gate)
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
        acc = fmaT(rv, rv, acc);
    }
    block_partial<BS>(acc, ra.partial, blockIdx.x);
}

// x += alpha*p if an iteration's x update is still pending (sc->xpend)
template <int BS>
__global__ __launch_bounds__(BS) void k_cg_xflush(int64_t n, const double* __restrict__ p, double* __restrict__ x,
                                                  const CGScalars* __restrict__ sc, const int* gate)
{
    if(gated(gate) || sc->xpend == 0) return;
    const double alpha = sc->alpha;
    int64_t b, e;
    chunk_range(n, b, e);
    for(int64_t i = b + threadIdx.x; i < e; i += BS) x[i] = fmaT(alpha, p[i], x[i]);
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

// max_i |x_i - y_i| (ComputeResidual.cpp:60-119): one partial per block, then
// one block takes the maximum of the partials (the maximum does not depend on
// the order, so any grouping gives the same value)
template <int BS>
__global__ __launch_bounds__(BS) void k_absdiff_max(int64_t n, const double* __restrict__ x,
                                                    const double* __restrict__ y, double* __restrict__ partial)
{
    int64_t b, e;
    chunk_range(n, b, e);
    double v = 0.0;
    for(int64_t i = b + threadIdx.x; i < e; i += BS) v = fmax(v, fabs(x[i] - y[i]));
#pragma unroll
    for(int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    __shared__ double lds[BS / 64];
    if((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
    __syncthreads();
    if(threadIdx.x == 0)
    {
        double mx = 0.0;
#pragma unroll
        for(int w = 0; w < BS / 64; ++w) mx = fmax(mx, lds[w]);
        partial[blockIdx.x] = mx;
    }
}

template <int BS>
__global__ __launch_bounds__(BS) void k_fold_max(int64_t n, const double* __restrict__ partial, double* __restrict__ out)
{
    double v = 0.0;
    for(int64_t i = threadIdx.x; i < n; i += BS) v = fmax(v, partial[i]);
#pragma unroll
    for(int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    __shared__ double lds[BS / 64];
    if((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
    __syncthreads();
    if(threadIdx.x == 0)
    {
        double mx = 0.0;
#pragma unroll
        for(int w = 0; w < BS / 64; ++w) mx = fmax(mx, lds[w]);
        *out = mx;
    }
}

// Diagonal access for TestCG (SparseMatrix.cpp:66-152)
template <typename VT>
__global__ void k_get_diag(int64_t m, const int32_t* __restrict__ dgi, const VT* __restrict__ val, double* __restrict__ d)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if(r < m) d[r] = cvt<double>(val[ell_at(dgi[r], r)]);
}

__global__ void k_set_diag(int64_t m, const int32_t* __restrict__ dgi, double* __restrict__ val,
                           double* __restrict__ dg, double* __restrict__ inv_diag, const double* __restrict__ d)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if(r < m)
    {
        val[ell_at(dgi[r], r)] = d[r];
        dg[r] = d[r];
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

// the fold that completes a reduction whose producer ran `nparts` blocks
int launch_fold(int64_t nparts, const RedArgs& ra, hipStream_t s, const int* gate)
{
    hipLaunchKernelGGL((k_fold<FOLD_THREADS>), dim3(1), dim3(FOLD_THREADS), 0, s, nparts, ra, gate);
    HPCGX_KCHECK();
    return 0;
}

// Launch shapes compiled in: rows per thread {1,2} x block {64,128,256}.
template <typename F>
static int with_shape(const Shape& sh, F&& f)
{
    using std::integral_constant;
    switch(sh.rows * 1000 + sh.block)
    {
    case 1064: return f(integral_constant<int, 1>{}, integral_constant<int, 64>{});
    case 1128: return f(integral_constant<int, 1>{}, integral_constant<int, 128>{});
    case 1256: return f(integral_constant<int, 1>{}, integral_constant<int, 256>{});
    case 2064: return f(integral_constant<int, 2>{}, integral_constant<int, 64>{});
    case 2128: return f(integral_constant<int, 2>{}, integral_constant<int, 128>{});
    case 2256: return f(integral_constant<int, 2>{}, integral_constant<int, 256>{});
    default: break;
    }
    set_error("unsupported launch shape (rows " + std::to_string(sh.rows) + ", block " + std::to_string(sh.block) + ")");
    return -2;
}

bool shape_supported(int rows, int block)
{
    return (rows == 1 || rows == 2) && (block == 64 || block == 128 || block == 256);
}

// R rows per thread need every colour block to be R-aligned (a thread's R
// rows then share their ELL slice).
static Shape fit(const Level& L, Shape sh)
{
    bool ok = true;
    for(int c = 0; c < L.ncolors && ok; ++c) ok = (L.offsets[c] % sh.rows) == 0 && (L.sizes[c] % sh.rows) == 0;
    if(!ok) sh.rows = 1;
    return sh;
}

// With a fused dot, a block walks tpb tiles so that the partials stay few
// enough for one fold; `resident` (blocks of the kernel the device keeps in
// flight at once) sets the group size, one XCD's share.
static Tiles make_tiles(const Level& L, int rows_per_tile, bool dot, int64_t resident, int64_t cap,
                        int64_t& nblocks)
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
    t.jtiles = (int)ceil_div(mx, rows_per_tile);
    const int64_t ntiles = (int64_t)t.ncolors * t.jtiles;
    t.tpb = 1;
    while(dot && ceil_div(ntiles, t.tpb) > cap) ++t.tpb;
    nblocks = ceil_div(ntiles, t.tpb);
    t.gx = (int)std::max<int64_t>(1, resident / NUM_XCD);
    return t;
}

// Workgroups of `kernel` (BS threads) the device holds at once, queried once
// per kernel instantiation.
template <typename K>
static int64_t resident_blocks(K kernel, int bs, int64_t& cache)
{
    if(cache > 0) return cache;
    int dev = 0, ncu = 0, per_cu = 0;
    if(hipGetDevice(&dev) == hipSuccess &&
       hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
       hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, bs, 0) == hipSuccess && ncu > 0 && per_cu > 0)
        cache = (int64_t)ncu * per_cu;
    else
        cache = 256 * 2;   // one MI355X at two blocks per CU
    return cache;
}

// element-type views of a level's arrays
template <typename VT>
static const VT* vals_of(const Level& L);
template <>
const double* vals_of<double>(const Level& L) { return L.ell_val; }
template <>
const __half* vals_of<__half>(const Level& L) { return L.ell_h; }

template <typename VT>
static const VT* diag_of(const Level& L);
template <>
const double* diag_of<double>(const Level& L) { return L.diag64; }
template <>
const __half* diag_of<__half>(const Level& L) { return L.diag_h; }

template <typename VT, typename XT, typename YT>
static int spmv_t(const Level& L, const void* x, void* y, const RedArgs* ra, hipStream_t s, const int* gate)
{
    const Shape sh = fit(L, L.shp[std::is_same<VT, double>::value ? PREC_F64 : PREC_MIXED][SH_SPMV]);
    return with_shape(sh, [&](auto RR, auto BB) -> int {
        constexpr int R = decltype(RR)::value, BS = decltype(BB)::value;
        int64_t nb;
        if(ra)
        {
            const int64_t cap = std::min<int64_t>(ra->cap, FOLD_MAX);
            static int64_t resident = 0;
            Tiles tl = make_tiles(L, BS * R, true, resident_blocks(k_spmv<VT, XT, YT, true, R, BS>, BS, resident),
                                  cap, nb);
            hipLaunchKernelGGL((k_spmv<VT, XT, YT, true, R, BS>), dim3((unsigned)nb), dim3(BS), 0, s, L.ell_col,
                               vals_of<VT>(L), (const XT*)x, (YT*)y, tl, *ra, gate);
            HPCGX_KCHECK();
            return launch_fold(nb, *ra, s, gate);
        }
        Tiles tl = make_tiles(L, BS * R, false, 0, 0, nb);
        hipLaunchKernelGGL((k_spmv<VT, XT, YT, false, R, BS>), dim3((unsigned)nb), dim3(BS), 0, s, L.ell_col,
                           vals_of<VT>(L), (const XT*)x, (YT*)y, tl, RedArgs{}, gate);
        HPCGX_KCHECK();
        return 0;
    });
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

// The sweeps' block order for a colour of `size` rows in tiles of `tile` rows
// (grid = its tile count).  Planes of the half-resolution colour grid whose x
// gathers (7 other colours, 3 planes, 8 B) outgrow an XCD's 4 MB L2 are cut
// into columns of >= min_rows rows (measured best: 8192 for the f64 sweeps,
// 4096 for the f16/f32 ones), so that the blocks in flight on an XCD march
// through one column's planes; Level::zcols > 0 forces a column count.
static ZMap zmap_for(const Level& L, int64_t size, int tile, unsigned grid, int64_t min_rows)
{
    ZMap z;
    if(L.zcols == 1 || L.ncolors != 8 || ((L.nx | L.ny | L.nz) & 1) != 0) return z;
    const int64_t plane = (int64_t)(L.nx / 2) * (L.ny / 2), planes = L.nz / 2;
    if(size != plane * planes || plane % tile) return z;
    const int64_t tpp = plane / tile;
    int64_t k = L.zcols;
    if(k == 0)
    {
        if(plane * 8 * 21 <= ((int64_t)6 << 20)) return z;
        k = 1;
        while(tpp % (2 * k) == 0 && (tpp / (2 * k)) * tile >= min_rows) k *= 2;
    }
    if(k <= 1 || tpp % k || (int64_t)grid != planes * tpp) return z;
    z.k = (int)k;
    z.tpc = (int)(tpp / k);
    z.planes = (int)planes;
    return z;
}

// one launch per colour c in [c0, c1) (step +1) or (c0, c1] (step -1); a
// colour whose slot window (kind `wk`, one row per thread) exceeds 16 slots
// runs with the op's wide shape, the others with its narrow shape
template <typename Launch>
static int per_color(const Level& L, int prec, int which, int wk, int c_first, int c_last, int step,
                     Launch&& launch)
{
    for(int c = c_first; step > 0 ? c <= c_last : c >= c_last; c += step)
    {
        if(L.sizes[c] == 0) continue;
        const bool wide = (wk < 0) || L.win[wk][c] > 16;
        const Shape sh = fit(L, L.shp[prec][wide ? which : which + SH_NARROW]);
        HPCGX_TRY(with_shape(sh, [&](auto RR, auto BB) -> int {
            constexpr int R = decltype(RR)::value, BS = decltype(BB)::value;
            const dim3 grid(nblk(ceil_div(L.sizes[c], R), BS));
            // odd colours walk their blocks backwards (in both sweeps: forward
            // ends on colour 7 reversed, backward starts on colour 6 forwards,
            // so every launch starts on the slab the previous one ended on)
            const int64_t size = (c & 1) ? -(int64_t)L.sizes[c] : (int64_t)L.sizes[c];
            launch(RR, BB, grid, c, size, zmap_for(L, L.sizes[c], BS * R, grid.x, prec == PREC_F64 ? 8192 : 4096));
            HPCGX_KCHECK();
            return 0;
        }));
    }
    return 0;
}

// The sweep kernels of a level: part 0 = zero-guess forward (colours 0..n-1),
// 1 = zero-guess backward (n-2..0), 2 = forward keeping t (0..n-1), 3 =
// backward from t (n-2..0), 4 = part 1 also keeping the upper sums outside
// the last colour in t, 5 = part 2 reusing them (valid after part 4 when
// only last-colour rows changed in between).
template <typename VT, typename XT, typename RT>
static int sweep_t(const Level& L, int part, const void* rv, void* xv, void* tv, hipStream_t s, const int* gate)
{
    const RT* r = (const RT*)rv;
    XT* x = (XT*)xv;
    XT* t = (XT*)tv;
    const VT* val = vals_of<VT>(L);
    const VT* dg = diag_of<VT>(L);
    const int last = L.ncolors - 1;
    const int pr = std::is_same<VT, double>::value ? PREC_F64 : PREC_MIXED;
    using std::integral_constant;
    switch(part)
    {
    case 0:
        return per_color(L, pr, SH_FWD_ZERO, WIN_LOWER, 0, last, 1, [&](auto RR, auto BB, dim3 grid, int c, int64_t sz, ZMap zm) {
            constexpr int R = decltype(RR)::value, BS = decltype(BB)::value;
            const int n = L.win[WIN_LOWER][c];
            with_window(n, [&](auto NN) {
                constexpr int NS = decltype(NN)::value;
                hipLaunchKernelGGL((k_fwd_zero<NS, VT, XT, RT, R, BS>), grid, dim3(BS), 0, s,
                                   (int64_t)L.offsets[c], sz, zm, L.ell_col, val, dg, L.desc, r, x, gate);
            });
        });
    case 1:
    case 4:
        if(last < 1) return 0;
        return per_color(L, pr, SH_BWD_ZERO, WIN_UPPER, last - 1, 0, -1, [&](auto RR, auto BB, dim3 grid, int c, int64_t sz, ZMap zm) {
            constexpr int R = decltype(RR)::value, BS = decltype(BB)::value;
            const int n = L.win[WIN_UPPER][c];
            with_window(n, [&](auto NN) {
                constexpr int NS = decltype(NN)::value;
                if(part == 4)
                    hipLaunchKernelGGL((k_bwd_zero<NS, VT, XT, true, R, BS>), grid, dim3(BS), 0, s,
                                       (int64_t)L.offsets[c], sz, zm, L.ell_col, val, dg, L.desc, x, t, gate);
                else
                    hipLaunchKernelGGL((k_bwd_zero<NS, VT, XT, false, R, BS>), grid, dim3(BS), 0, s,
                                       (int64_t)L.offsets[c], sz, zm, L.ell_col, val, dg, L.desc, x, t, gate);
            });
        });
    case 2:
    case 5:
        return per_color(L, pr, SH_FWD_FULL, part == 5 ? WIN_LOWER : -1, 0, last, 1, [&](auto RR, auto BB, dim3 grid, int c, int64_t sz, ZMap zm) {
            constexpr int R = decltype(RR)::value, BS = decltype(BB)::value;
            if(c < last && part == 5)
            {
                const int na = L.win[WIN_LOWER][c], nb = L.win[WIN_TAIL][c];
                with_window_pair(na, nb, [&](auto NA_, auto NB_) {
                    constexpr int NA = decltype(NA_)::value, NB = decltype(NB_)::value;
                    hipLaunchKernelGGL((k_fwd_full_u<NA, NB, VT, XT, RT, R, BS>), grid, dim3(BS), 0, s,
                                       (int64_t)L.offsets[c], sz, zm, L.ell_col, val, dg, L.desc, r, x, t,
                                       gate);
                });
            }
            else if(c < last)
                hipLaunchKernelGGL((k_fwd_full<VT, XT, RT, true, R, BS>), grid, dim3(BS), 0, s,
                                   (int64_t)L.offsets[c], sz, zm, L.ell_col, val, r, x, t, gate);
            else
                hipLaunchKernelGGL((k_fwd_full<VT, XT, RT, false, R, BS>), grid, dim3(BS), 0, s,
                                   (int64_t)L.offsets[c], sz, zm, L.ell_col, val, r, x, t, gate);
        });
    case 3:
        if(last < 1) return 0;
        return per_color(L, pr, SH_BWD_FULL, WIN_UPPER, last - 1, 0, -1, [&](auto RR, auto BB, dim3 grid, int c, int64_t sz, ZMap zm) {
            constexpr int R = decltype(RR)::value, BS = decltype(BB)::value;
            const int n = L.win[WIN_UPPER][c];
            with_window(n, [&](auto NN) {
                constexpr int NS = decltype(NN)::value;
                hipLaunchKernelGGL((k_bwd_full<NS, VT, XT, R, BS>), grid, dim3(BS), 0, s,
                                   (int64_t)L.offsets[c], sz, zm, L.ell_col, val, dg, L.desc, (const XT*)t, x, gate);
            });
        });
    }
    set_error("invalid sweep part");
    return -2;
}

int window_slots(int w)
{
    int ns = 0;
    with_window(w, [&](auto NN) { ns = decltype(NN)::value; });
    return ns;
}

void window_pair_slots(int wl, int wt, int* a, int* b)
{
    with_window_pair(wl, wt, [&](auto NA_, auto NB_) {
        *a = decltype(NA_)::value;
        *b = decltype(NB_)::value;
    });
}

int launch_sweep(const Level& L, Prec prec, int part, const void* r, bool r_f64, void* x, void* t, hipStream_t s,
                 const int* gate)
{
    if(prec == PREC_F64) return sweep_t<double, double, double>(L, part, r, x, t, s, gate);
    if(r_f64) return sweep_t<__half, float, double>(L, part, r, x, t, s, gate);
    return sweep_t<__half, float, float>(L, part, r, x, t, s, gate);
}

int launch_symgs_zero(const Level& L, Prec prec, const void* r, bool r_f64, void* x, hipStream_t s, const int* gate,
                      void* u)
{
    HPCGX_TRY(launch_sweep(L, prec, 0, r, r_f64, x, nullptr, s, gate));
    return launch_sweep(L, prec, u ? 4 : 1, r, r_f64, x, u, s, gate);
}

int launch_symgs(const Level& L, Prec prec, const void* r, bool r_f64, void* x, void* t, hipStream_t s,
                 const int* gate, bool use_u)
{
    HPCGX_TRY(launch_sweep(L, prec, use_u ? 5 : 2, r, r_f64, x, t, s, gate));
    return launch_sweep(L, prec, 3, r, r_f64, x, t, s, gate);
}

template <typename VT, typename XT, typename RT>
static int restrict_t(const Level& F, const void* r, const void* x, void* rc, hipStream_t s, const int* gate)
{
    const int bs = F.shp[std::is_same<VT, double>::value ? PREC_F64 : PREC_MIXED][SH_RESTRICT].block;
    auto go = [&](auto BB) -> int {
        constexpr int BS = decltype(BB)::value;
        hipLaunchKernelGGL((k_restrict<VT, XT, RT, XT, BS>), dim3(nblk(F.mc, BS)), dim3(BS), 0, s, F.mc,
                           F.ell_col, vals_of<VT>(F), F.f2c_fine, F.f2c_coarse, (const RT*)r, (const XT*)x,
                           (XT*)rc, gate);
        HPCGX_KCHECK();
        return 0;
    };
    if(bs == 64) return go(std::integral_constant<int, 64>{});
    if(bs == 128) return go(std::integral_constant<int, 128>{});
    return go(std::integral_constant<int, 256>{});
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

int launch_spmv_injected(const Level& F, const double* x, double* y, hipStream_t s)
{
    if(!F.ell_val || F.mc == 0)
    {
        set_error("injected-row SpMV needs an f64 level with a coarser level");
        return -2;
    }
    hipLaunchKernelGGL(k_spmv_rows, dim3(nblk(F.mc, 256)), dim3(256), 0, s, F.mc, F.ell_col, F.ell_val, F.f2c_fine,
                       x, y);
    HPCGX_KCHECK();
    return 0;
}

int launch_restrict_diff(const Level& F, const double* rf, const double* Axf, double* rc, hipStream_t s)
{
    hipLaunchKernelGGL(k_restrict_diff, dim3(nblk(F.mc, 256)), dim3(256), 0, s, F.mc, F.f2c_fine, F.f2c_coarse, rf,
                       Axf, rc);
    HPCGX_KCHECK();
    return 0;
}

// The read ceiling of a level's matrix stream: every byte of its column and
// value arrays read once with 16-byte non-temporal loads, four in flight per
// thread, and no gathers -- what the SpMV streams, without the rest of its
// work.  The xor of the data decides a store that never happens in practice,
// so the loads cannot be dropped.
typedef uint32_t U32x4 __attribute__((ext_vector_type(4)));

template <int BS>
__global__ __launch_bounds__(BS) void k_read_stream(int64_t na, const U32x4* __restrict__ a, int64_t nb,
                                                    const U32x4* __restrict__ b, double* __restrict__ sink)
{
    const int64_t stride = (int64_t)gridDim.x * BS;
    U32x4 acc = {0u, 0u, 0u, 0u};
    auto pass = [&](int64_t n, const U32x4* __restrict__ v) {
        int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
        for(; i + 3 * stride < n; i += 4 * stride)
        {
            const U32x4 w0 = __builtin_nontemporal_load(v + i);
            const U32x4 w1 = __builtin_nontemporal_load(v + i + stride);
            const U32x4 w2 = __builtin_nontemporal_load(v + i + 2 * stride);
            const U32x4 w3 = __builtin_nontemporal_load(v + i + 3 * stride);
            acc ^= (w0 ^ w1) ^ (w2 ^ w3);
        }
        for(; i < n; i += stride) acc ^= __builtin_nontemporal_load(v + i);
    };
    pass(na, a);
    pass(nb, b);
    if((acc.x ^ acc.y ^ acc.z ^ acc.w) == 0x9e3779b9u) sink[blockIdx.x] = 1.0;
}

int launch_read_stream(const Level& L, Prec prec, double* sink, int64_t sink_cap, hipStream_t s)
{
    const int64_t me = ell_rows(L.m);
    const void* val = prec == PREC_F64 ? (const void*)L.ell_val : (const void*)L.ell_h;
    const int64_t vbytes = W * me * (prec == PREC_F64 ? 8 : 2);
    const int64_t cbytes = W * me * 4;
    if(!val || (vbytes % 16) || (cbytes % 16))
    {
        set_error("read stream needs the level's value store and 16-byte multiples");
        return -2;
    }
    constexpr int BS = 256;
    const unsigned g = (unsigned)std::min<int64_t>(RED_BLOCKS, sink_cap);
    hipLaunchKernelGGL((k_read_stream<BS>), dim3(g), dim3(BS), 0, s, cbytes / 16, (const U32x4*)L.ell_col,
                       vbytes / 16, (const U32x4*)val, sink);
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

int64_t reduction_blocks(int64_t n) { return red_grid(n); }

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
    return launch_fold(g, ra, s, gate);
}

int launch_absdiff_max(int64_t n, const double* x, const double* y, double* partial, int64_t cap, double* out,
                       hipStream_t s)
{
    constexpr int BS = RED_THREADS;
    const unsigned g = (unsigned)std::min<int64_t>(red_grid(n), cap);
    hipLaunchKernelGGL((k_absdiff_max<BS>), dim3(g), dim3(BS), 0, s, n, x, y, partial);
    HPCGX_KCHECK();
    hipLaunchKernelGGL((k_fold_max<FOLD_THREADS>), dim3(1), dim3(FOLD_THREADS), 0, s, (int64_t)g,
                       (const double*)partial, out);
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
    const unsigned g = red_grid(n);
    hipLaunchKernelGGL((k_axpy_dot<BS>), dim3(g), dim3(BS), 0, s, n, alpha, x, y, ra, gate);
    HPCGX_KCHECK();
    return launch_fold(g, ra, s, gate);
}

int launch_cg_p(int64_t n, const void* z, bool z_f64, double* p, double* x, const CGScalars* sc, hipStream_t s,
                const int* gate)
{
    constexpr int BS = RED_THREADS;
    if(z_f64)
        hipLaunchKernelGGL((k_cg_p<double, BS>), dim3(red_grid(n)), dim3(BS), 0, s, n, (const double*)z, p, x, sc,
                           gate);
    else
        hipLaunchKernelGGL((k_cg_p<float, BS>), dim3(red_grid(n)), dim3(BS), 0, s, n, (const float*)z, p, x, sc,
                           gate);
    HPCGX_KCHECK();
    return 0;
}

int launch_cg_xr(int64_t n, const double* Ap, double* r, const RedArgs& ra, hipStream_t s, const int* gate)
{
    constexpr int BS = RED_THREADS;
    const unsigned g = red_grid(n);
    hipLaunchKernelGGL((k_cg_xr<BS>), dim3(g), dim3(BS), 0, s, n, Ap, r, ra, gate);
    HPCGX_KCHECK();
    return launch_fold(g, ra, s, gate);
}

int launch_cg_xflush(int64_t n, const double* p, double* x, const CGScalars* sc, hipStream_t s, const int* gate)
{
    constexpr int BS = RED_THREADS;
    hipLaunchKernelGGL((k_cg_xflush<BS>), dim3(red_grid(n)), dim3(BS), 0, s, n, p, x, sc, gate);
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
    hipLaunchKernelGGL(k_set_diag, dim3(nblk(L.m, 256)), dim3(256), 0, s, L.m, L.diag_idx, L.ell_val, L.diag64, L.inv_diag, d);
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
