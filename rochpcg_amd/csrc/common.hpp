// common.hpp -- shared types and device helpers of the MI355X (gfx950) HPCG
// solver library.  Everything here is internal; the public surfaces are
// include/rochpcg_amd.h (C ABI) and include/hpcg/*.hpp (reference-named C++).
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

namespace hpcgx
{

constexpr int W = 27;            // ELL width of the 27-point operator
constexpr int MAX_COLORS = 128;  // colours a level may hold (the reference's MAX_COLORS, utils.hpp:48)
constexpr int MAX_LEVELS = 8;
constexpr int NUM_XCD = 8;       // XCDs of an MI355X

// Error reporting: every entry point returns an int status; the last message
// is kept per thread for hpcgx_last_error().
void set_error(const std::string& msg);
const char* last_error();

#define HPCGX_HIP(expr)                                                                       \
    do                                                                                         \
    {                                                                                          \
        hipError_t e_ = (expr);                                                                \
        if(e_ != hipSuccess)                                                                   \
        {                                                                                      \
            ::hpcgx::set_error(std::string(#expr) + ": " + hipGetErrorString(e_) + " at " +   \
                               __FILE__ + ":" + std::to_string(__LINE__));                     \
            return -1;                                                                         \
        }                                                                                      \
    } while(0)

#define HPCGX_TRY(expr)                                                                        \
    do                                                                                         \
    {                                                                                          \
        int s_ = (expr);                                                                       \
        if(s_ != 0) return s_;                                                                 \
    } while(0)

// Precision of a level's arrays.
enum Prec : int
{
    PREC_F64 = 0,   // f64 ELL values, f64 vectors
    PREC_MIXED = 1  // f16 ELL values, f32 vectors
};

// Launch shape of one operator on one level (chosen by the autotuner).
struct Shape
{
    int block = 256;   // threads per workgroup: 64, 128 or 256
    int rows = 1;      // consecutive rows per thread: 1 or 2
};

// Block-to-tile order of one colour launch of a sweep.  k <= 1: tile = block.
// k > 1: the colour's tiles form `planes` planes of k columns of tpc tiles
// each (in row order); block b takes column b / (planes*tpc), plane
// (b / tpc) % planes, tile b % tpc of that column -- so the blocks in flight on
// an XCD march through the planes of one column.  Any such map is a bijection
// of the colour's tiles, and the rows of a colour are independent, so results
// do not depend on it.
struct ZMap
{
    int k = 1;
    int tpc = 0;      // tiles per column and plane
    int planes = 0;
};

// Slot windows of the sweeps (see load_window): per colour, the widest window
// of its rows in the split-padded ELL (the same for every rows-per-thread R,
// since each window starts at a fixed slot)
enum WinKind : int
{
    WIN_LOWER = 0,   // [0, d): max d
    WIN_UPPER = 1,   // [us, W): max W - us (upper entries)
    WIN_TAIL = 2,    // [tl, W): max W - tl (last-colour entries)
    WIN_COUNT = 3
};

// tunable operators of a level.  A sweep has two shapes: one for its wide
// colours (slot window over 16) and one (SH_NARROW + op) for the others.
enum ShapeOp : int
{
    SH_SPMV = 0,
    SH_FWD_ZERO = 1,
    SH_BWD_ZERO = 2,
    SH_FWD_FULL = 3,
    SH_BWD_FULL = 4,
    SH_RESTRICT = 5,
    SH_FWD_ZERO_N = 6,
    SH_BWD_ZERO_N = 7,
    SH_FWD_FULL_N = 8,
    SH_BWD_FULL_N = 9,
    SH_COUNT = 10
};
constexpr int SH_NARROW = SH_FWD_ZERO_N - SH_FWD_ZERO;

// One level of the hierarchy, device pointers plus host-side colour table.
// Rows are in multicolour order: colour c owns rows [offsets[c], offsets[c+1]).
struct Level
{
    int nx = 0, ny = 0, nz = 0;
    int64_t m = 0;          // rows
    int64_t nnz = 0;        // stored nonzeros
    int ncolors = 0;
    int sizes[MAX_COLORS] = {};
    int offsets[MAX_COLORS + 1] = {};
    int round_sizes[MAX_COLORS] = {};   // per JPL round, as the reference records them

    // sliced ELL: entry p of row r lives at ell_at(p, r) (64-bit offsets); the
    // entries of a row in ascending column order, split-padded (kernels.hpp)
    int32_t* ell_col = nullptr;   // [W*ell_rows(m)], -1 padding between the diagonal and the right-aligned upper entries
    double* ell_val = nullptr;    // [W*ell_rows(m)] (f64 levels)
    __half* ell_h = nullptr;      // [W*ell_rows(m)] (mixed levels)
    int32_t* diag_idx = nullptr;  // [m] slot of the diagonal
    int32_t* desc = nullptr;      // [m] row descriptor: diagonal slot, first last-colour slot, first upper slot
    double* inv_diag = nullptr;   // [m] 1/a_ii (f64)
    double* diag64 = nullptr;     // [m] a_ii (f64 levels)
    __half* diag_h = nullptr;     // [m] a_ii (mixed levels)
    int32_t* perm = nullptr;      // [m] natural row -> multicolour row
    int32_t* colors = nullptr;    // [m] colour of each natural row
    int32_t* iperm_tmp = nullptr; // [m] multicolour row -> natural row (setup only)

    // transfer to the next coarser level (indexed by coarse natural row k)
    int64_t mc = 0;
    int32_t* f2c_fine = nullptr;    // [mc] multicolour fine row injected into k
    int32_t* f2c_coarse = nullptr;  // [mc] multicolour coarse row of k
    int32_t* f2c_nat = nullptr;     // [mc] natural fine row (reference f2cOperator)
    // every injected fine row is in the last colour: the prolongation changes
    // only last-colour rows, so the post-smoother may reuse the pre-smoother's
    // upper sums over the other colours
    bool u_ok = false;
    unsigned char win[WIN_COUNT][MAX_COLORS] = {};   // [kind][colour]: slots a sweep thread loads per row
    // columns per plane of the sweeps' block order (ZMap): 0 chooses, 1 keeps
    // the row order
    int zcols = 0;

    Shape shp[2][SH_COUNT];   // [precision][operator]
};

// Device-side CG scalars (one cache line each to avoid false sharing is not
// needed: a single block writes them).
struct CGScalars
{
    double rtz;
    double oldrtz;
    double pAp;
    double alpha;
    double beta;
    double normr;
    double normr0;
    double tol;
    int iter;       // iterations completed
    int done;       // 1 once normr/normr0 <= tol or max_iter reached
    int max_iter;
    int xpend;      // single-problem CG: x still lacks the last iteration's alpha*p
};

// Deterministic two-phase reductions: every block of the producing kernel
// writes one partial, then a one-block fold kernel sums them in a fixed order.
constexpr int RED_BLOCKS = 4096;    // blocks of a streaming vector reduction
constexpr int RED_THREADS = 256;
constexpr int FOLD_MAX = 16384;     // partials a single-problem fold reads at most
constexpr int FOLD_THREADS = 1024;

// ---------------------------------------------------------------------------
// device helpers

template <typename T>
__device__ __forceinline__ T ntload(const T* p)
{
    return __builtin_nontemporal_load(p);
}
__device__ __forceinline__ __half ntload(const __half* p)
{
    return __ushort_as_half(__builtin_nontemporal_load(reinterpret_cast<const unsigned short*>(p)));
}

// Sliced ELL.  Rows are grouped in slices of ELL_SLICE consecutive rows;
// slice s stores slot p of its rows contiguously at [(s*W + p)*ELL_SLICE,
// +ELL_SLICE), so entry p of row r lives at ell_at(p, r).  A wave's slot loads
// (64 or 128 rows of one slot) stay one contiguous segment, as in a
// column-major ELL, but the 27 slots of a slice are neighbours in memory: a
// wave streams one 5 KB (f16) to 41 KB (int32 + f64) block per array instead of
// 27 segments 4*rows bytes apart, which HBM serves at a higher rate.  The
// arrays hold ell_rows(m) = m rounded up to a whole slice; the rows past m are
// padding (-1, 0).
constexpr int64_t ELL_SLICE = 128;

__host__ __device__ __forceinline__ int64_t ell_at(int p, int64_t r)
{
    const uint64_t u = (uint64_t)r;   // rows are non-negative: shifts, not divisions
    return (int64_t)(((u / ELL_SLICE) * W + (uint64_t)p) * ELL_SLICE + (u % ELL_SLICE));
}

__host__ __device__ __forceinline__ int64_t ell_rows(int64_t m) { return (m + ELL_SLICE - 1) / ELL_SLICE * ELL_SLICE; }

// Contiguous block ranges per XCD: the hardware hands workgroup i to XCD
// i % 8, so logical block L = base(i % 8) + i / 8 gives each XCD one
// contiguous slice of the row space (its L2 then serves the x-gathers of
// neighbouring rows).
__host__ __device__ __forceinline__ int64_t xcd_block(int64_t i, int64_t nblocks)
{
    const int64_t q = nblocks / NUM_XCD;
    const int64_t rr = nblocks % NUM_XCD;
    const int64_t x = i % NUM_XCD;
    const int64_t base = x * q + (x < rr ? x : rr);
    return base + i / NUM_XCD;
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for(int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Block sum of BS threads (BS/64 waves), result valid in thread 0.
template <int BS>
__device__ __forceinline__ double block_sum(double v, double* lds)
{
    v = wave_sum(v);
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    if(lane == 0) lds[wid] = v;
    __syncthreads();
    double s = 0.0;
    if(threadIdx.x == 0)
    {
#pragma unroll
        for(int w = 0; w < BS / 64; ++w) s += lds[w];
    }
    return s;
}

// ---------------------------------------------------------------------------
// host helpers
inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// A device allocation freed when it goes out of scope (every early error
// return included) unless release() hands it on.
struct DevTemp
{
    void* p = nullptr;
    DevTemp() = default;
    DevTemp(const DevTemp&) = delete;
    DevTemp& operator=(const DevTemp&) = delete;
    ~DevTemp()
    {
        if(p) (void)hipFree(p);
    }
    void* release()
    {
        void* q = p;
        p = nullptr;
        return q;
    }
};

}  // namespace hpcgx
