// setup.hip -- device-side problem setup for one level of the HPCG hierarchy:
// JPL multicolouring, the colour-stable permutation, the multicolour-ordered
// column-major ELL operator, the injection maps to the next coarser level and
// the right-hand sides.  Everything is built on the GPU from the grid
// dimensions; no natural-order matrix is ever materialised.
//
// Reference behaviour restated here:
//   stencil / values / b       GenerateProblem.cpp:98-303, GenerateProblem_ref.cpp:131-182
//   row hash                   GenerateProblem.cpp:93-96, :294-295
//   JPL rounds and colours     MultiColoring.cpp:151-236 (kernel_jpl), :238-327
//   stable sort into perm      MultiColoring.cpp:328-344
//   column map + ascending     Permute.cpp:95-149 (kernel_perm_cols)
//   row permutation / ELL      Permute.cpp:54-76, :187-218; SparseMatrix.cpp:166-226
//   diagonal slot, 1/a_ii      SparseMatrix.cpp:361-389
//   injection operator         GenerateCoarseProblem.cpp:57-86
#include "common.hpp"
#include "setup.hpp"

#include <algorithm>
#include <climits>
#include <cstdlib>
#include <cstring>

namespace hpcgx
{

struct Grid
{
    int nx, ny, nz;
    __device__ __forceinline__ void coords(int64_t i, int& ix, int& iy, int& iz) const
    {
        const int64_t nxy = (int64_t)nx * ny;
        iz = (int)(i / nxy);
        const int64_t rem = i - (int64_t)iz * nxy;
        iy = (int)(rem / nx);
        ix = (int)(rem - (int64_t)iy * nx);
    }
    __device__ __forceinline__ int64_t hash(int ix, int iy, int iz) const
    {
        const int64_t cls = ((ix & 1) << 2) | ((iy & 1) << 1) | (iz & 1);
        const int64_t nxy = (int64_t)nx * ny;
        return cls * nxy * nz + (int64_t)iz * nxy + (int64_t)iy * (nx << 1) + (ix << 2);
    }
};

__global__ void k_fill_i32(int64_t n, int32_t* p, int32_t v)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if(i < n) p[i] = v;
}

// One JPL round over natural rows.  The decision reads the colours before the
// round (cin) and writes cout, which is what the reference's in-place kernel
// computes: rows coloured in the same round hold color1/color2 and are treated
// as uncoloured by its neighbour test.
__global__ __launch_bounds__(256) void k_jpl_round(Grid g, int64_t m, const int32_t* __restrict__ cin,
                                                   int32_t* __restrict__ cout, int color1, int color2,
                                                   unsigned long long* __restrict__ counts)
{
    const int64_t row = (int64_t)blockIdx.x * 256 + threadIdx.x;
    int out = -1;
    if(row < m)
    {
        const int cur = cin[row];
        out = cur;
        if(cur == -1)
        {
            int ix, iy, iz;
            g.coords(row, ix, iy, iz);
            const int64_t rh = g.hash(ix, iy, iz);
            bool is_max = true, is_min = true;
            for(int sz = -1; sz <= 1; ++sz)
            {
                const int jz = iz + sz;
                if(jz < 0 || jz >= g.nz) continue;
                for(int sy = -1; sy <= 1; ++sy)
                {
                    const int jy = iy + sy;
                    if(jy < 0 || jy >= g.ny) continue;
                    for(int sx = -1; sx <= 1; ++sx)
                    {
                        const int jx = ix + sx;
                        if(jx < 0 || jx >= g.nx) continue;
                        if(sx == 0 && sy == 0 && sz == 0) continue;
                        const int64_t j = ((int64_t)jz * g.ny + jy) * g.nx + jx;
                        const int cn = cin[j];
                        if(cn == -1 || cn == color1 || cn == color2)
                        {
                            const int64_t jh = g.hash(jx, jy, jz);
                            // equal hashes (small odd levels such as 4x3x2) never
                            // resolve under the reference's >= / <= tests; the row
                            // index breaks the tie, leaving every other case unchanged
                            if(jh > rh || (jh == rh && j > row)) is_max = false;
                            if(jh < rh || (jh == rh && j < row)) is_min = false;
                        }
                    }
                }
            }
            if(is_max)
                out = color1;
            else if(is_min)
                out = color2;
        }
        cout[row] = out;
    }
    // count rows holding color1 / color2 (the reference counts over all rows)
    const unsigned long long b1 = __ballot(row < m && out == color1);
    const unsigned long long b2 = __ballot(row < m && out == color2);
    if((threadIdx.x & 63) == 0)
    {
        if(b1) atomicAdd(counts + 0, (unsigned long long)__popcll(b1));
        if(b2) atomicAdd(counts + 1, (unsigned long long)__popcll(b2));
    }
}

// Per-block colour histogram (block = 256 natural rows).
__global__ __launch_bounds__(256) void k_color_hist(int64_t m, int ncolors, const int32_t* __restrict__ colors,
                                                    int32_t* __restrict__ hist, int64_t nblk)
{
    __shared__ int cnt[4][MAX_COLORS];
    const int64_t row = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int c = (row < m) ? colors[row] : -1;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for(int cc = 0; cc < ncolors; ++cc)
    {
        const unsigned long long b = __ballot(c == cc);
        if(lane == 0) cnt[wid][cc] = __popcll(b);
    }
    __syncthreads();
    if((int)threadIdx.x < ncolors)
    {
        const int cc = threadIdx.x;
        hist[(int64_t)cc * nblk + blockIdx.x] = cnt[0][cc] + cnt[1][cc] + cnt[2][cc] + cnt[3][cc];
    }
}

// perm[i] = base[colour][block] + rank of i among the block's rows of that
// colour (stable: natural order is kept within each colour).
__global__ __launch_bounds__(256) void k_color_scatter(int64_t m, int ncolors, const int32_t* __restrict__ colors,
                                                       const int64_t* __restrict__ base, int64_t nblk,
                                                       int32_t* __restrict__ perm, int32_t* __restrict__ iperm)
{
    __shared__ int cnt[4][MAX_COLORS];
    const int64_t row = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int c = (row < m) ? colors[row] : -1;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    int rank = 0;
    for(int cc = 0; cc < ncolors; ++cc)
    {
        const unsigned long long b = __ballot(c == cc);
        if(lane == 0) cnt[wid][cc] = __popcll(b);
        if(c == cc) rank = __popcll(b & lt);
    }
    __syncthreads();
    if(row < m)
    {
        for(int w = 0; w < wid; ++w) rank += cnt[w][c];
        const int64_t p = base[(int64_t)c * nblk + blockIdx.x] + rank;
        perm[row] = (int32_t)p;
        iperm[p] = (int32_t)row;
    }
}

template <typename T>
__device__ __forceinline__ void cswap(T& a, T& b)
{
    const T lo = a < b ? a : b;
    const T hi = a < b ? b : a;
    a = lo;
    b = hi;
}

// Build multicolour row r of the column-major ELL: its natural row i =
// iperm[r]; the (up to) 27 neighbours are mapped through perm and sorted
// ascending (odd-even transposition network over statically indexed
// registers).  Split padding: lower entries and the diagonal from slot 0, the
// upper entries right-aligned to slot W-1, padding (-1, 0.0) in between; the
// descriptor gives the diagonal slot, the first last-colour slot and the
// first upper slot (W when absent).
template <typename VT>
__global__ __launch_bounds__(256) void k_build_ell(Grid g, int64_t m, const int32_t* __restrict__ perm,
                                                   const int32_t* __restrict__ iperm,
                                                   int32_t* __restrict__ ell_col, VT* __restrict__ ell_val,
                                                   int32_t* __restrict__ diag_idx, int32_t* __restrict__ desc,
                                                   int32_t last_off, double* __restrict__ inv_diag,
                                                   VT* __restrict__ dvals)
{
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if(r >= ell_rows(m)) return;
    if(r >= m)
    {
        // padding rows of the last slice
#pragma unroll
        for(int p = 0; p < W; ++p)
        {
            ell_col[ell_at(p, r)] = -1;
            ell_val[ell_at(p, r)] = (VT)0.0f;
        }
        return;
    }
    const int64_t i = iperm[r];
    int ix, iy, iz;
    g.coords(i, ix, iy, iz);
    int key[W];
#pragma unroll
    for(int s = 0; s < W; ++s)
    {
        const int sz = s / 9 - 1, sy = (s / 3) % 3 - 1, sx = s % 3 - 1;
        const int jx = ix + sx, jy = iy + sy, jz = iz + sz;
        const bool in = jx >= 0 && jx < g.nx && jy >= 0 && jy < g.ny && jz >= 0 && jz < g.nz;
        key[s] = in ? perm[((int64_t)jz * g.ny + jy) * g.nx + jx] : INT_MAX;
    }
#pragma unroll
    for(int round = 0; round < W; ++round)
    {
#pragma unroll
        for(int k = (round & 1); k + 1 < W; k += 2) cswap(key[k], key[k + 1]);
    }
    // sorted: [0, ne) valid, the diagonal at dg, the off-diagonal last-colour
    // columns from tl0
    int dg = 0, ne = 0, tl0 = W;
#pragma unroll
    for(int p = 0; p < W; ++p)
    {
        const int c = key[p];
        if(c != INT_MAX) ne = p + 1;
        if((int64_t)c == r) dg = p;
        if(c != INT_MAX && c >= last_off && (int64_t)c != r && p < tl0) tl0 = p;
    }
    if(tl0 > ne) tl0 = ne;
    // split padding: shift the upper entries right by W - ne (a barrel shift
    // over statically indexed registers), leaving [dg+1, dg+1+shift) empty
    const int shift = W - ne;
#pragma unroll
    for(int k = 16; k >= 1; k >>= 1)
    {
        const bool on = (shift & k) != 0;
#pragma unroll
        for(int p = W - 1; p >= 0; --p)
        {
            const int v = (p - k >= 0 && p - k > dg) ? key[p - k >= 0 ? p - k : 0] : (p > dg ? INT_MAX : key[p]);
            if(on) key[p] = v;
        }
    }
#pragma unroll
    for(int p = 0; p < W; ++p)
    {
        const int c = key[p];
        const bool valid = (c != INT_MAX);
        ell_col[ell_at(p, r)] = valid ? c : -1;
        ell_val[ell_at(p, r)] = (VT)(valid ? (p == dg ? 26.0f : -1.0f) : 0.0f);
    }
    diag_idx[r] = dg;
    desc[r] = dg | ((tl0 + shift) << 8) | ((dg + 1 + shift) << 16);
    if(inv_diag) inv_diag[r] = 1.0 / 26.0;
    dvals[r] = (VT)26.0f;
}

// Injection maps for coarse natural row k: f2c_nat = 2*kc per dimension
// (GenerateCoarseProblem.cpp:77-84), plus the multicolour rows on both sides.
__global__ __launch_bounds__(256) void k_transfer(Grid gc, int nxf, int nyf, int64_t mc,
                                                  const int32_t* __restrict__ perm_f,
                                                  const int32_t* __restrict__ perm_c, int32_t* __restrict__ f2c_nat,
                                                  int32_t* __restrict__ f2c_fine, int32_t* __restrict__ f2c_coarse)
{
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if(k >= mc) return;
    int kx, ky, kz;
    gc.coords(k, kx, ky, kz);
    const int64_t f = (int64_t)(2 * kz) * nxf * nyf + (int64_t)(2 * ky) * nxf + 2 * kx;
    f2c_nat[k] = (int32_t)f;
    f2c_fine[k] = perm_f[f];
    f2c_coarse[k] = perm_c[k];
}

// splitmix64 output i of `seed`, mapped to [0,1) with 53 bits.
__device__ __forceinline__ double uniform01(uint64_t seed, int64_t i)
{
    uint64_t z = seed + (uint64_t)(i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

// Right-hand sides in multicolour order (natural order when perm is null).  kind 0: xexact = 1, b = 27 - nnz_row
// (GenerateProblem.cpp:300); kind 1: xexact = 2u-1, b = A xexact summed in
// natural stencil order with fma.
__global__ __launch_bounds__(256) void k_rhs(Grid g, int64_t m, int kind, uint64_t seed,
                                             const int32_t* __restrict__ perm, double* __restrict__ b,
                                             double* __restrict__ xexact, double* __restrict__ x0)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if(i >= m) return;
    int ix, iy, iz;
    g.coords(i, ix, iy, iz);
    double bv, xv;
    if(kind == 0)
    {
        const int k = (1 + (ix > 0) + (ix < g.nx - 1)) * (1 + (iy > 0) + (iy < g.ny - 1)) *
                      (1 + (iz > 0) + (iz < g.nz - 1));
        bv = 26.0 - ((double)(k - 1));
        xv = 1.0;
    }
    else
    {
        double sum = 0.0;
        for(int sz = -1; sz <= 1; ++sz)
        {
            const int jz = iz + sz;
            if(jz < 0 || jz >= g.nz) continue;
            for(int sy = -1; sy <= 1; ++sy)
            {
                const int jy = iy + sy;
                if(jy < 0 || jy >= g.ny) continue;
                for(int sx = -1; sx <= 1; ++sx)
                {
                    const int jx = ix + sx;
                    if(jx < 0 || jx >= g.nx) continue;
                    const int64_t j = ((int64_t)jz * g.ny + jy) * g.nx + jx;
                    const double a = (j == i) ? 26.0 : -1.0;
                    sum = fma(a, 2.0 * uniform01(seed, j) - 1.0, sum);
                }
            }
        }
        bv = sum;
        xv = 2.0 * uniform01(seed, i) - 1.0;
    }
    const int64_t p = perm ? perm[i] : i;
    if(b) b[p] = bv;
    if(xexact) xexact[p] = xv;
    if(x0) x0[p] = 0.0;
}

// Natural-order injection maps: f2c[k] = fine row of coarse row k, c2f[f] =
// coarse row of fine row f (the fill kernel sets -1 first).
__global__ __launch_bounds__(256) void k_f2c_natural(Grid gc, int nxf, int nyf, int64_t mc,
                                                     int32_t* __restrict__ f2c, int32_t* __restrict__ c2f)
{
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if(k >= mc) return;
    int kx, ky, kz;
    gc.coords(k, kx, ky, kz);
    const int64_t f = (int64_t)(2 * kz) * nxf * nyf + (int64_t)(2 * ky) * nxf + 2 * kx;
    f2c[k] = (int32_t)f;
    if(c2f) c2f[f] = (int32_t)k;
}

// out[perm[i]] = in[i]
__global__ __launch_bounds__(256) void k_permute(int64_t m, const int32_t* __restrict__ perm, const double* __restrict__ in,
                                                 double* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if(i < m) out[perm[i]] = in[i];
}

// ---------------------------------------------------------------------------
static inline unsigned grid_of(int64_t n, int bs) { return (unsigned)ceil_div(n, bs); }

int natural_transfer(int nxf, int nyf, int nzf, int32_t* f2c, int32_t* c2f, hipStream_t s)
{
    Grid gc{nxf / 2, nyf / 2, nzf / 2};
    const int64_t mc = (int64_t)gc.nx * gc.ny * gc.nz;
    const int64_t mf = (int64_t)nxf * nyf * nzf;
    if(c2f) hipLaunchKernelGGL(k_fill_i32, dim3(grid_of(mf, 256)), dim3(256), 0, s, mf, c2f, -1);
    hipLaunchKernelGGL(k_f2c_natural, dim3(grid_of(mc, 256)), dim3(256), 0, s, gc, nxf, nyf, mc, f2c, c2f);
    HPCGX_HIP(hipGetLastError());
    return 0;
}

int permute_vector(int64_t m, const int32_t* perm, const double* in, double* out, hipStream_t s)
{
    hipLaunchKernelGGL(k_permute, dim3(grid_of(m, 256)), dim3(256), 0, s, m, perm, in, out);
    HPCGX_HIP(hipGetLastError());
    return 0;
}

__global__ __launch_bounds__(256) void k_iota_i64(int64_t n, int64_t* __restrict__ p)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if(i < n) p[i] = i;
}

int iota_i64(int64_t n, int64_t* p, hipStream_t s)
{
    if(n <= 0) return 0;
    hipLaunchKernelGGL(k_iota_i64, dim3(grid_of(n, 256)), dim3(256), 0, s, n, p);
    HPCGX_HIP(hipGetLastError());
    return 0;
}

int natural_rhs(int nx, int ny, int nz, double* b, double* xexact, double* x0, hipStream_t s)
{
    Grid g{nx, ny, nz};
    const int64_t m = (int64_t)nx * ny * nz;
    hipLaunchKernelGGL(k_rhs, dim3(grid_of(m, 256)), dim3(256), 0, s, g, m, 0, (uint64_t)0, (const int32_t*)nullptr, b,
                       xexact, x0);
    HPCGX_HIP(hipGetLastError());
    return 0;
}

int color_level(Level& L, hipStream_t s)
{
    const int64_t m = L.m;
    Grid g{L.nx, L.ny, L.nz};
    // device temporaries, freed on every return; the final colours and the
    // permutations are handed to the level (release) once complete
    DevTemp ta, tb, tcounts, thist, tbase, tperm, tiperm;
    HPCGX_HIP(hipMalloc(&ta.p, sizeof(int32_t) * m));
    HPCGX_HIP(hipMalloc(&tb.p, sizeof(int32_t) * m));
    HPCGX_HIP(hipMalloc(&tcounts.p, 2 * sizeof(unsigned long long)));
    unsigned long long* counts = (unsigned long long*)tcounts.p;
    hipLaunchKernelGGL(k_fill_i32, dim3(grid_of(m, 256)), dim3(256), 0, s, m, (int32_t*)ta.p, -1);
    // Colour draws as in the reference: srand(0x586744), rand() % 8 for the
    // first eight colours, then counting upwards.
    srand(0x586744);
    int nblocks = 0;
    int64_t colored = 0;
    int rounds = 0;
    while(colored != m)
    {
        if(nblocks + 2 > MAX_COLORS)
        {
            set_error("JPL colouring needs more than MAX_COLORS colours");
            return -2;
        }
        const int c1 = (nblocks < 8) ? rand() % 8 : nblocks;
        const int c2 = (nblocks < 8) ? rand() % 8 : nblocks + 1;
        HPCGX_HIP(hipMemsetAsync(counts, 0, 2 * sizeof(unsigned long long), s));
        hipLaunchKernelGGL(k_jpl_round, dim3(grid_of(m, 256)), dim3(256), 0, s, g, m, (const int32_t*)ta.p,
                           (int32_t*)tb.p, c1, c2, counts);
        HPCGX_HIP(hipGetLastError());
        unsigned long long h[2];
        HPCGX_HIP(hipMemcpyAsync(h, counts, sizeof(h), hipMemcpyDeviceToHost, s));
        HPCGX_HIP(hipStreamSynchronize(s));
        L.round_sizes[nblocks] = (int)h[0];
        L.round_sizes[nblocks + 1] = (int)h[1];
        colored += (int64_t)h[0] + (int64_t)h[1];
        nblocks += 2;
        std::swap(ta.p, tb.p);
        if(++rounds > 4 * MAX_COLORS) break;
    }
    if(colored != m)
    {
        set_error("JPL colouring did not colour every row");
        return -2;
    }
    const int32_t* colors = (const int32_t*)ta.p;

    // per-colour counts via the block histogram; ncolors = highest colour + 1.
    // The colours drawn are below max(8, nblocks): rand() % 8 in the first
    // rounds, then nblocks and nblocks + 1.
    const int ncb = std::max(8, nblocks);
    const int64_t nblk = ceil_div(m, 256);
    HPCGX_HIP(hipMalloc(&thist.p, sizeof(int32_t) * nblk * ncb));
    HPCGX_HIP(hipMalloc(&tbase.p, sizeof(int64_t) * nblk * ncb));
    int32_t* hist = (int32_t*)thist.p;
    int64_t* base = (int64_t*)tbase.p;
    hipLaunchKernelGGL(k_color_hist, dim3((unsigned)nblk), dim3(256), 0, s, m, ncb, colors, hist, nblk);
    HPCGX_HIP(hipGetLastError());
    std::vector<int32_t> hh((size_t)nblk * ncb);
    HPCGX_HIP(hipMemcpyAsync(hh.data(), hist, sizeof(int32_t) * hh.size(), hipMemcpyDeviceToHost, s));
    HPCGX_HIP(hipStreamSynchronize(s));
    int64_t tot[MAX_COLORS] = {};
    int ncolors = 0;
    for(int c = 0; c < ncb; ++c)
    {
        for(int64_t b = 0; b < nblk; ++b) tot[c] += hh[(size_t)c * nblk + b];
        if(tot[c] > 0) ncolors = c + 1;
    }
    L.ncolors = ncolors;
    std::vector<int64_t> hb((size_t)nblk * ncb);
    int64_t run = 0;
    for(int c = 0; c < ncolors; ++c)
    {
        L.offsets[c] = (int)run;
        L.sizes[c] = (int)tot[c];
        for(int64_t b = 0; b < nblk; ++b)
        {
            hb[(size_t)c * nblk + b] = run;
            run += hh[(size_t)c * nblk + b];
        }
    }
    L.offsets[ncolors] = (int)run;
    HPCGX_HIP(hipMemcpyAsync(base, hb.data(), sizeof(int64_t) * hb.size(), hipMemcpyHostToDevice, s));
    HPCGX_HIP(hipMalloc(&tperm.p, sizeof(int32_t) * m));
    HPCGX_HIP(hipMalloc(&tiperm.p, sizeof(int32_t) * m));
    hipLaunchKernelGGL(k_color_scatter, dim3((unsigned)nblk), dim3(256), 0, s, m, ncolors, colors, base,
                       nblk, (int32_t*)tperm.p, (int32_t*)tiperm.p);
    HPCGX_HIP(hipGetLastError());
    HPCGX_HIP(hipStreamSynchronize(s));
    L.colors = (int32_t*)ta.release();
    L.perm = (int32_t*)tperm.release();
    L.iperm_tmp = (int32_t*)tiperm.release();
    return 0;
}

// Window statistics of one colour's rows [off, off+size): the largest lower
// count d, upper count W - us and last-colour count W - tl (stats[0..2]).
__global__ __launch_bounds__(256) void k_window_stats(int64_t off, int64_t size, const int32_t* __restrict__ desc,
                                                      int* __restrict__ stats)
{
    int mx[WIN_COUNT] = {0, 0, 0};
    for(int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < size; i += (int64_t)gridDim.x * 256)
    {
        const int v = desc[off + i];
        mx[WIN_LOWER] = max(mx[WIN_LOWER], v & 255);
        mx[WIN_TAIL] = max(mx[WIN_TAIL], W - ((v >> 8) & 255));
        mx[WIN_UPPER] = max(mx[WIN_UPPER], W - ((v >> 16) & 255));
    }
    __shared__ int lmx[WIN_COUNT][4];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for(int k = 0; k < WIN_COUNT; ++k)
    {
        int v = mx[k];
        for(int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
        if(lane == 0) lmx[k][wid] = v;
    }
    __syncthreads();
    if(threadIdx.x == 0)
        for(int k = 0; k < WIN_COUNT; ++k)
            atomicMax(stats + k, max(max(lmx[k][0], lmx[k][1]), max(lmx[k][2], lmx[k][3])));
}

// Per colour, the slots a sweep thread loads per row of each window kind
// (Level::win), from the row descriptors.
static int window_stats(Level& L, hipStream_t s)
{
    const int nc = L.ncolors;
    int* stats = nullptr;
    HPCGX_HIP(hipMalloc(&stats, sizeof(int) * WIN_COUNT * nc));
    HPCGX_HIP(hipMemsetAsync(stats, 0, sizeof(int) * WIN_COUNT * nc, s));
    for(int c = 0; c < nc; ++c)
    {
        if(L.sizes[c] == 0) continue;
        const int64_t g = std::min<int64_t>(1024, ceil_div(L.sizes[c], 256));
        hipLaunchKernelGGL(k_window_stats, dim3((unsigned)g), dim3(256), 0, s, (int64_t)L.offsets[c],
                           (int64_t)L.sizes[c], L.desc, stats + WIN_COUNT * c);
        HPCGX_HIP(hipGetLastError());
    }
    std::vector<int> hs((size_t)WIN_COUNT * nc);
    HPCGX_HIP(hipMemcpyAsync(hs.data(), stats, sizeof(int) * WIN_COUNT * nc, hipMemcpyDeviceToHost, s));
    HPCGX_HIP(hipStreamSynchronize(s));
    HPCGX_HIP(hipFree(stats));
    for(int c = 0; c < nc; ++c)
        for(int k = 0; k < WIN_COUNT; ++k) L.win[k][c] = (unsigned char)hs[WIN_COUNT * c + k];
    return 0;
}

int build_level(Level& L, int nx, int ny, int nz, bool with_f64, bool with_f16, hipStream_t s)
{
    L.nx = nx;
    L.ny = ny;
    L.nz = nz;
    L.m = (int64_t)nx * ny * nz;
    if(L.m <= 0 || L.m > INT_MAX)
    {
        set_error("level size out of range");
        return -2;
    }
    L.nnz = (int64_t)(3 * nx - 2) * (3 * ny - 2) * (3 * nz - 2);
    HPCGX_TRY(color_level(L, s));
    const int64_t m = L.m;
    Grid g{nx, ny, nz};
    HPCGX_HIP(hipMalloc(&L.ell_col, sizeof(int32_t) * W * ell_rows(m)));
    HPCGX_HIP(hipMalloc(&L.diag_idx, sizeof(int32_t) * m));
    HPCGX_HIP(hipMalloc(&L.desc, sizeof(int32_t) * m));
    const int32_t last_off = L.offsets[L.ncolors - 1];
    if(with_f64)
    {
        HPCGX_HIP(hipMalloc(&L.ell_val, sizeof(double) * W * ell_rows(m)));
        HPCGX_HIP(hipMalloc(&L.inv_diag, sizeof(double) * m));
        HPCGX_HIP(hipMalloc(&L.diag64, sizeof(double) * m));
        hipLaunchKernelGGL(k_build_ell<double>, dim3(grid_of(ell_rows(m), 256)), dim3(256), 0, s, g, m, L.perm,
                           L.iperm_tmp, L.ell_col, L.ell_val, L.diag_idx, L.desc, last_off, L.inv_diag,
                           L.diag64);
        HPCGX_HIP(hipGetLastError());
    }
    if(with_f16)
    {
        HPCGX_HIP(hipMalloc(&L.ell_h, sizeof(__half) * W * ell_rows(m)));
        HPCGX_HIP(hipMalloc(&L.diag_h, sizeof(__half) * m));
        hipLaunchKernelGGL(k_build_ell<__half>, dim3(grid_of(ell_rows(m), 256)), dim3(256), 0, s, g, m, L.perm,
                           L.iperm_tmp, L.ell_col, L.ell_h, L.diag_idx, L.desc, last_off, (double*)nullptr,
                           L.diag_h);
        HPCGX_HIP(hipGetLastError());
    }
    HPCGX_HIP(hipStreamSynchronize(s));
    HPCGX_HIP(hipFree(L.iperm_tmp));
    L.iperm_tmp = nullptr;
    return window_stats(L, s);
}

int build_transfer(Level& F, Level& C, hipStream_t s)
{
    F.mc = C.m;
    Grid gc{C.nx, C.ny, C.nz};
    HPCGX_HIP(hipMalloc(&F.f2c_nat, sizeof(int32_t) * F.mc));
    HPCGX_HIP(hipMalloc(&F.f2c_fine, sizeof(int32_t) * F.mc));
    HPCGX_HIP(hipMalloc(&F.f2c_coarse, sizeof(int32_t) * F.mc));
    hipLaunchKernelGGL(k_transfer, dim3(grid_of(F.mc, 256)), dim3(256), 0, s, gc, F.nx, F.ny, F.mc, F.perm,
                       C.perm, F.f2c_nat, F.f2c_fine, F.f2c_coarse);
    HPCGX_HIP(hipGetLastError());
    // are all injected rows in the last colour? (true for the parity-class
    // colourings JPL produces on these grids; checked, not assumed)
    std::vector<int32_t> h((size_t)F.mc);
    HPCGX_HIP(hipMemcpyAsync(h.data(), F.f2c_fine, sizeof(int32_t) * F.mc, hipMemcpyDeviceToHost, s));
    HPCGX_HIP(hipStreamSynchronize(s));
    const int32_t last_off = F.offsets[F.ncolors - 1];
    F.u_ok = F.ncolors > 1;
    for(int32_t v : h) F.u_ok = F.u_ok && v >= last_off;
    return 0;
}

int make_rhs(const Level& L, int kind, uint64_t seed, double* b, double* xexact, double* x0, hipStream_t s)
{
    Grid g{L.nx, L.ny, L.nz};
    hipLaunchKernelGGL(k_rhs, dim3(grid_of(L.m, 256)), dim3(256), 0, s, g, L.m, kind, seed, L.perm, b, xexact,
                       x0);
    HPCGX_HIP(hipGetLastError());
    return 0;
}

void free_level(Level& L)
{
    auto f = [](void* p) {
        if(p) (void)hipFree(p);
    };
    f(L.ell_col);
    f(L.ell_val);
    f(L.ell_h);
    f(L.diag_idx);
    f(L.desc);
    f(L.inv_diag);
    f(L.diag64);
    f(L.diag_h);
    f(L.perm);
    f(L.colors);
    f(L.iperm_tmp);
    f(L.f2c_fine);
    f(L.f2c_coarse);
    f(L.f2c_nat);
    L = Level();
}

}  // namespace hpcgx
