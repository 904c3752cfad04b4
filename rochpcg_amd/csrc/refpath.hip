// refpath.hip -- the reference's natural-order CPU path on the device.
//
// The benchmark driver takes its reference tolerance from 50 iterations of
// CG_ref in natural row order (main.cpp:334-344), before the matrix is
// coloured.  This file runs that path on the GPU with the reference's own
// arithmetic:
//   SpMV          ComputeSPMV_ref.cpp:47-72     y_i = sum_j a_ij x_j, ascending j
//   SymGS         ComputeSYMGS_ref.cpp:54-103   sequential Gauss-Seidel, forward
//                                               then backward, natural order
//   V-cycle       ComputeMG_ref.cpp:39-63       x = 0, SymGS, A x, injection,
//                                               recursion, prolongation, SymGS
//   restriction   ComputeRestriction_ref.cpp:40-54
//   prolongation  ComputeProlongation_ref.cpp:38-52
//   WAXPBY        ComputeWAXPBY_ref.cpp:42-70
//   CG            CG_ref.cpp:86-174
//
// The operator is the 27-point stencil of GenerateProblem_ref.cpp:131-182
// (diagonal 26, neighbours -1), applied from the grid coordinates: the
// entries of natural row (ix, iy, iz) in ascending column order are the
// in-range offsets sz, sy, sx in {-1, 0, 1}, nested z -> y -> x.  Products and
// sums are rounded one by one (no fused multiply-add), in the reference's
// order, so SpMV, SymGS, the V-cycle and the vector updates give the
// reference's bits; only the dot products (a fixed two-phase tree here, an
// OpenMP reduction there) round differently.
//
// Sequential Gauss-Seidel as wavefronts.  Row (ix, iy, iz) belongs to
// wavefront w = ix + 2 iy + 4 iz.  A neighbour at offset (dx, dy, dz) lies in
// wavefront w + dx + 2 dy + 4 dz; every neighbour with a smaller natural index
// (dz = -1, or dz = 0 and dy = -1, or dz = dy = 0 and dx = -1) has a smaller w,
// every neighbour with a larger index a larger w, and no two neighbours share
// a wavefront (dx + 2 dy + 4 dz = 0 forces dx = dy = dz = 0).  Sweeping the
// wavefronts in increasing order, each one launch, therefore reads exactly
// the values the sequential forward sweep reads: lower neighbours already
// updated, upper neighbours not yet.  The backward sweep runs the wavefronts
// in decreasing order.
#include "refpath.hpp"

#include <algorithm>
#include <cstring>

namespace hpcgx
{

struct NatGrid
{
    int nx, ny, nz;
};

// sum_j a_ij x_j over natural row (ix, iy, iz), ascending j, rounded per
// product and per sum (ComputeSPMV_ref.cpp:62-68)
__device__ __forceinline__ double ref_row_sum(const NatGrid& g, int ix, int iy, int iz, const double* __restrict__ x)
{
#pragma clang fp contract(off)
    const int64_t nxy = (int64_t)g.nx * g.ny;
    const int64_t i = iz * nxy + (int64_t)iy * g.nx + ix;
    double sum = 0.0;
    for(int sz = -1; sz <= 1; ++sz)
    {
        const int z = iz + sz;
        if(z < 0 || z >= g.nz) continue;
        for(int sy = -1; sy <= 1; ++sy)
        {
            const int y = iy + sy;
            if(y < 0 || y >= g.ny) continue;
            for(int sx = -1; sx <= 1; ++sx)
            {
                const int xx = ix + sx;
                if(xx < 0 || xx >= g.nx) continue;
                const int64_t j = z * nxy + (int64_t)y * g.nx + xx;
                const double a = (j == i) ? 26.0 : -1.0;
                sum += a * x[j];
            }
        }
    }
    return sum;
}

__global__ __launch_bounds__(256) void k_ref_spmv(NatGrid g, int64_t m, const double* __restrict__ x,
                                                  double* __restrict__ y)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if(i >= m) return;
    const int64_t nxy = (int64_t)g.nx * g.ny;
    const int iz = (int)(i / nxy);
    const int64_t rem = i - iz * nxy;
    const int iy = (int)(rem / g.nx);
    const int ix = (int)(rem - (int64_t)iy * g.nx);
    y[i] = ref_row_sum(g, ix, iy, iz, x);
}

// One wavefront w of a Gauss-Seidel sweep (ComputeSYMGS_ref.cpp:67-100).
// Thread t covers iz = zlo + t / ylen and the (t % ylen)-th iy of that plane
// with 0 <= w - 2 iy - 4 iz < nx; ix = w - 2 iy - 4 iz.
__global__ __launch_bounds__(256) void k_ref_gs_wave(NatGrid g, int w, int zlo, int ylen, int64_t nthreads,
                                                     const double* __restrict__ r, double* __restrict__ x)
{
#pragma clang fp contract(off)
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if(t >= nthreads) return;
    const int iz = zlo + (int)(t / ylen);
    const int u = w - 4 * iz;           // ix + 2 iy, >= 0 since iz <= w / 4
    const int v = u - (g.nx - 1);       // 2 iy >= v keeps ix < nx
    const int iy = (v > 0 ? (v + 1) / 2 : 0) + (int)(t % ylen);
    if(iy > g.ny - 1 || 2 * iy > u) return;
    const int ix = u - 2 * iy;
    const int64_t i = iz * ((int64_t)g.nx * g.ny) + (int64_t)iy * g.nx + ix;
    // sum = r_i - sum_j a_ij x_j (the diagonal included), then the diagonal
    // term added back, as the reference writes it
    double sum = r[i];
    const int64_t nxy = (int64_t)g.nx * g.ny;
    for(int sz = -1; sz <= 1; ++sz)
    {
        const int z = iz + sz;
        if(z < 0 || z >= g.nz) continue;
        for(int sy = -1; sy <= 1; ++sy)
        {
            const int y = iy + sy;
            if(y < 0 || y >= g.ny) continue;
            for(int sx = -1; sx <= 1; ++sx)
            {
                const int xx = ix + sx;
                if(xx < 0 || xx >= g.nx) continue;
                const int64_t j = z * nxy + (int64_t)y * g.nx + xx;
                const double a = (j == i) ? 26.0 : -1.0;
                sum -= a * x[j];
            }
        }
    }
    sum += x[i] * 26.0;
    x[i] = sum / 26.0;
}

// rc[k] = rf[f] - (A xf)[f], f = f2c[k] (ComputeMG_ref.cpp:50-52 computes the
// whole A xf; the restriction reads it at the injected rows only)
__global__ __launch_bounds__(256) void k_ref_restrict(NatGrid gf, int64_t mc, const int32_t* __restrict__ f2c,
                                                      const double* __restrict__ rf, const double* __restrict__ xf,
                                                      double* __restrict__ rc)
{
#pragma clang fp contract(off)
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if(k >= mc) return;
    const int64_t f = f2c[k];
    const int64_t nxy = (int64_t)gf.nx * gf.ny;
    const int iz = (int)(f / nxy);
    const int64_t rem = f - iz * nxy;
    const int iy = (int)(rem / gf.nx);
    const int ix = (int)(rem - (int64_t)iy * gf.nx);
    const double axf = ref_row_sum(gf, ix, iy, iz, xf);
    rc[k] = rf[f] - axf;
}

__global__ __launch_bounds__(256) void k_ref_prolong(int64_t mc, const int32_t* __restrict__ f2c,
                                                     const double* __restrict__ xc, double* __restrict__ xf)
{
#pragma clang fp contract(off)
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if(k >= mc) return;
    const int64_t f = f2c[k];
    xf[f] += xc[k];
}

// w = x + beta y (the alpha == 1 branch of ComputeWAXPBY_ref.cpp:53-56), beta
// given or read on the device: sel 0 = beta, 1 = +alpha, 2 = -alpha, 3 = the
// CG direction update (p = z on the first iteration, else z + beta p)
__global__ __launch_bounds__(256) void k_ref_xpby(int64_t n, const double* __restrict__ x, double beta,
                                                  const double* __restrict__ y, double* __restrict__ w,
                                                  const CGScalars* __restrict__ sc, int sel)
{
#pragma clang fp contract(off)
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if(i >= n) return;
    double b = beta;
    if(sel == 1) b = sc->alpha;
    if(sel == 2) b = -sc->alpha;
    if(sel == 3)
    {
        if(sc->iter == 0)
        {
            w[i] = x[i];
            return;
        }
        b = sc->beta;
    }
    w[i] = x[i] + b * y[i];
}

__global__ __launch_bounds__(256) void k_ref_fill(int64_t n, double* __restrict__ x, double v)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if(i < n) x[i] = v;
}

__global__ __launch_bounds__(256) void k_ref_copy(int64_t n, const double* __restrict__ x, double* __restrict__ y)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if(i < n) y[i] = x[i];
}

// out[i] = in[perm[i]]: a multicolour vector taken back to natural order
__global__ __launch_bounds__(256) void k_ref_unpermute(int64_t m, const int32_t* __restrict__ perm,
                                                       const double* __restrict__ in, double* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if(i < m) out[i] = in[perm[i]];
}

// ---------------------------------------------------------------------------
static inline unsigned blocks_of(int64_t n) { return (unsigned)ceil_div(n, 256); }

static NatGrid grid_of(const Level& L) { return NatGrid{L.nx, L.ny, L.nz}; }

void refpath_free(Problem& P)
{
    RefPath* R = P.ref;
    if(!R) return;
    auto f = [](void* q) {
        if(q) (void)hipFree(q);
    };
    f(R->b);
    f(R->x);
    f(R->r);
    f(R->z);
    f(R->p);
    f(R->Ap);
    for(int l = 0; l < MAX_LEVELS; ++l)
    {
        f(R->lr[l]);
        f(R->lx[l]);
    }
    f(R->sc);
    f(R->hist);
    delete R;
    P.ref = nullptr;
}

static int refpath_ensure(Problem& P)
{
    if(P.ref) return 0;
    RefPath* R = new RefPath;
    P.ref = R;
    const size_t m0 = (size_t)P.L[0].m;
    double** v0[] = {&R->b, &R->x, &R->r, &R->z, &R->p, &R->Ap};
    int rc = 0;
    for(double** v : v0)
        if(!rc && hipMalloc((void**)v, m0 * 8) != hipSuccess) rc = -1;
    for(int l = 1; l < P.nlevels && !rc; ++l)
    {
        const size_t m = (size_t)P.L[l].m;
        if(hipMalloc((void**)&R->lr[l], m * 8) != hipSuccess || hipMalloc((void**)&R->lx[l], m * 8) != hipSuccess)
            rc = -1;
    }
    R->hist_cap = P.hist_cap;
    if(!rc && (hipMalloc((void**)&R->sc, sizeof(CGScalars)) != hipSuccess ||
               hipMalloc((void**)&R->hist, sizeof(double) * (R->hist_cap + 1)) != hipSuccess))
        rc = -1;
    if(rc)
    {
        refpath_free(P);
        set_error("out of device memory for the natural-order reference path");
    }
    return rc;
}

int ref_spmv(Problem& P, int l, const double* x, double* y)
{
    const Level& L = P.L[l];
    hipLaunchKernelGGL(k_ref_spmv, dim3(blocks_of(L.m)), dim3(256), 0, P.s, grid_of(L), L.m, x, y);
    HPCGX_HIP(hipGetLastError());
    return 0;
}

// one sweep: wavefronts 0..wmax (forward) or wmax..0 (backward)
static int ref_sweep(Problem& P, int l, const double* r, double* x, bool forward)
{
    const Level& L = P.L[l];
    const NatGrid g = grid_of(L);
    const int wmax = (g.nx - 1) + 2 * (g.ny - 1) + 4 * (g.nz - 1);
    const int ylen = g.nx / 2 + 1;
    for(int k = 0; k <= wmax; ++k)
    {
        const int w = forward ? k : wmax - k;
        // planes holding a row of this wavefront: 0 <= w - 4 iz <= (nx-1) + 2 (ny-1)
        const int lo_num = w - (g.nx - 1) - 2 * (g.ny - 1);
        const int zlo = lo_num > 0 ? (lo_num + 3) / 4 : 0;
        const int zhi = std::min(g.nz - 1, w / 4);
        if(zhi < zlo) continue;
        const int64_t nthreads = (int64_t)(zhi - zlo + 1) * ylen;
        hipLaunchKernelGGL(k_ref_gs_wave, dim3(blocks_of(nthreads)), dim3(256), 0, P.s, g, w, zlo, ylen, nthreads, r,
                           x);
    }
    HPCGX_HIP(hipGetLastError());
    return 0;
}

int ref_symgs(Problem& P, int l, const double* r, double* x)
{
    HPCGX_TRY(ref_sweep(P, l, r, x, true));
    return ref_sweep(P, l, r, x, false);
}

int ref_mg(Problem& P, int l, const double* r, double* x)
{
    HPCGX_TRY(refpath_ensure(P));
    RefPath& R = *P.ref;
    const Level& L = P.L[l];
    hipLaunchKernelGGL(k_ref_fill, dim3(blocks_of(L.m)), dim3(256), 0, P.s, L.m, x, 0.0);
    HPCGX_HIP(hipGetLastError());
    HPCGX_TRY(ref_symgs(P, l, r, x));
    if(l + 1 >= P.nlevels) return 0;
    hipLaunchKernelGGL(k_ref_restrict, dim3(blocks_of(L.mc)), dim3(256), 0, P.s, grid_of(L), L.mc, L.f2c_nat, r, x,
                       R.lr[l + 1]);
    HPCGX_HIP(hipGetLastError());
    HPCGX_TRY(ref_mg(P, l + 1, R.lr[l + 1], R.lx[l + 1]));
    hipLaunchKernelGGL(k_ref_prolong, dim3(blocks_of(L.mc)), dim3(256), 0, P.s, L.mc, L.f2c_nat,
                       (const double*)R.lx[l + 1], x);
    HPCGX_HIP(hipGetLastError());
    return ref_symgs(P, l, r, x);
}

static int xpby(Problem& P, const double* x, double beta, const double* y, double* w, int sel)
{
    const int64_t n = P.L[0].m;
    hipLaunchKernelGGL(k_ref_xpby, dim3(blocks_of(n)), dim3(256), 0, P.s, n, x, beta, y, w,
                       (const CGScalars*)P.ref->sc, sel);
    HPCGX_HIP(hipGetLastError());
    return 0;
}

static int ref_dot(Problem& P, const double* x, const double* y, int op)
{
    RedArgs ra;
    ra.partial = P.partial;
    ra.op = op;
    ra.sc = P.ref->sc;
    ra.hist = P.ref->hist;
    ra.hist_cap = P.ref->hist_cap;
    ra.out = P.dout;
    ra.cap = P.partial_cap;
    return launch_dot(P.L[0].m, x, true, y, true, ra, P.s, nullptr);
}

int ref_cg(Problem& P, const double* b, double* x, int max_iter, double tol, bool precond, int* niters,
           double* normr, double* normr0, double* history, int hist_len)
{
    if(max_iter < 0 || max_iter > P.hist_cap)
    {
        set_error("max_iter out of range (history capacity " + std::to_string(P.hist_cap) + ")");
        return -2;
    }
    HPCGX_TRY(refpath_ensure(P));
    RefPath& R = *P.ref;
    const int64_t m = P.L[0].m;
    if(!b)
    {
        hipLaunchKernelGGL(k_ref_unpermute, dim3(blocks_of(m)), dim3(256), 0, P.s, m, (const int32_t*)P.L[0].perm,
                           (const double*)P.b, R.b);
        HPCGX_HIP(hipGetLastError());
        b = R.b;
    }
    if(!x)
    {
        hipLaunchKernelGGL(k_ref_fill, dim3(blocks_of(m)), dim3(256), 0, P.s, m, R.x, 0.0);
        HPCGX_HIP(hipGetLastError());
        x = R.x;
    }
    CGScalars h;
    std::memset(&h, 0, sizeof(h));
    h.tol = tol;
    h.max_iter = max_iter;
    HPCGX_HIP(hipMemcpyAsync(R.sc, &h, sizeof(h), hipMemcpyHostToDevice, P.s));
    // CG_ref.cpp:114-118: p = x; Ap = A p; r = b - Ap; normr = ||r||
    hipLaunchKernelGGL(k_ref_copy, dim3(blocks_of(m)), dim3(256), 0, P.s, m, (const double*)x, R.p);
    HPCGX_HIP(hipGetLastError());
    HPCGX_TRY(ref_spmv(P, 0, R.p, R.Ap));
    HPCGX_TRY(xpby(P, b, -1.0, R.Ap, R.r, 0));
    HPCGX_TRY(ref_dot(P, R.r, R.r, OP_NORMR0));
    CGScalars cur;
    for(int k = 1; k <= max_iter; ++k)
    {
        // the loop condition normr/normr0 > tolerance, as the fold left it
        HPCGX_HIP(hipMemcpyAsync(&cur, R.sc, sizeof(cur), hipMemcpyDeviceToHost, P.s));
        HPCGX_HIP(hipStreamSynchronize(P.s));
        if(cur.done) break;
        // CG_ref.cpp:135-160
        if(precond)
            HPCGX_TRY(ref_mg(P, 0, R.r, R.z));
        else
            HPCGX_TRY(xpby(P, R.r, 0.0, R.r, R.z, 0));
        HPCGX_TRY(ref_dot(P, R.r, R.z, OP_RTZ));   // beta = rtz / oldrtz after the first iteration
        HPCGX_TRY(xpby(P, R.z, 0.0, R.p, R.p, 3));  // p = z, then p = z + beta p
        HPCGX_TRY(ref_spmv(P, 0, R.p, R.Ap));
        HPCGX_TRY(ref_dot(P, R.p, R.Ap, OP_PAP));  // alpha = rtz / pAp
        HPCGX_TRY(xpby(P, x, 0.0, R.p, x, 1));     // x = x + alpha p
        HPCGX_TRY(xpby(P, R.r, 0.0, R.Ap, R.r, 2)); // r = r - alpha Ap
        HPCGX_TRY(ref_dot(P, R.r, R.r, OP_NORMR));
    }
    HPCGX_HIP(hipMemcpyAsync(&cur, R.sc, sizeof(cur), hipMemcpyDeviceToHost, P.s));
    HPCGX_HIP(hipStreamSynchronize(P.s));
    if(niters) *niters = cur.iter;
    if(normr) *normr = cur.normr;
    if(normr0) *normr0 = cur.normr0;
    if(history && hist_len > 0)
    {
        int n = std::min(cur.iter + 1, hist_len);
        HPCGX_HIP(hipMemcpy(history, R.hist, sizeof(double) * n, hipMemcpyDeviceToHost));
    }
    return 0;
}

}  // namespace hpcgx
