// batch.hpp -- many independent HPCG problems of uneven sizes solved together:
// each problem keeps its own hierarchy (a Problem), and every operator of
// the V-cycle and the CG runs as ONE launch per (operator, level, colour)
// across all problems.  Blocks find their (problem, row block) segment in a
// device table; reductions, CG scalars and convergence flags stay per problem.
#pragma once
#include "solver.hpp"

namespace hpcgx
{

// One contiguous row range of one problem in one launch.
struct BSeg
{
    int64_t blk0;      // first (logical) block of the segment in the launch
    int64_t nblk;      // blocks of the segment
    int64_t off;       // first row (or element) of the segment
    int64_t size;      // rows (or elements) in the segment
    const int32_t* col;
    const void* val;
    const int32_t* desc;   // row descriptors (Level::desc)
    const void* dg;    // diagonal values (Level::diag64 / diag_h)
    const void* r;     // rhs / first vector operand
    void* x;           // solution / in-out operand
    void* t;           // SymGS partial sums (null: do not keep them); upper sums in a
                       // pre-smoother backward sweep (null: not kept)
    int use_u;         // forward sweep reuses the upper sums in t (row_fwd_full_u)
    const int32_t* f2c_fine;
    const int32_t* f2c_coarse;
    void* rc;          // coarse vector (restriction output / prolongation input)
    const void* a;     // vector kernels: operands
    const void* b;
    void* w;
    const int* gate;   // the problem's convergence flag
    RedArgs ra;        // the problem's reduction scratch and scalars
};

// A launch's device table is its nseg BSeg records followed by the segment of
// every logical block, one uint16 each (map[b] = j for b in [blk0_j, blk0_j +
// nblk_j)), so a block finds its segment with one load instead of a search.
constexpr size_t MAX_SEGS = 65535;

__host__ __device__ inline const uint16_t* seg_map(const BSeg* segs, int nseg)
{
    return reinterpret_cast<const uint16_t*>(segs + nseg);
}

inline size_t table_bytes(size_t nseg, int64_t nblocks) { return sizeof(BSeg) * nseg + 2 * (size_t)nblocks; }

// fill the block map of segments whose blk0/nblk assign_blocks has set
inline void block_map(const BSeg* segs, size_t nseg, uint16_t* map)
{
    for(size_t j = 0; j < nseg; ++j)
        for(int64_t b = segs[j].blk0; b < segs[j].blk0 + segs[j].nblk; ++b) map[b] = (uint16_t)j;
}

// Blocks of the segments of one launch, in order: a vector launch gives each
// segment ceil(size / (4 block)) blocks, at most RED_BLOCKS and at least one;
// a row launch ceil(ceil(size / rows) / block).  Sets nblk and blk0 and
// returns the launch's total block count.
inline int64_t assign_blocks(BSeg* segs, size_t nseg, int block, bool vector_kind, int rows)
{
    int64_t nb = 0;
    for(size_t j = 0; j < nseg; ++j)
    {
        BSeg& sg = segs[j];
        if(vector_kind)
        {
            const int64_t g = ceil_div(sg.size, (int64_t)block * 4);
            sg.nblk = g < 1 ? 1 : (g > RED_BLOCKS ? RED_BLOCKS : g);
        }
        else
        {
            sg.nblk = ceil_div(ceil_div(sg.size, (int64_t)rows), (int64_t)block);
        }
        sg.blk0 = nb;
        nb += sg.nblk;
    }
    return nb;
}

// Rows per thread a sweep launch can use: 2 loads R consecutive rows of a
// slot as one vector, which needs even offsets and sizes in every segment
// (an even row's pair then shares its ELL slice); otherwise 1.
inline int launch_rows(const BSeg* segs, size_t nseg, int rows)
{
    for(size_t j = 0; j < nseg; ++j)
        if((segs[j].off % 2) != 0 || (segs[j].size % 2) != 0) return 1;
    return rows;
}

enum BKind : int
{
    BK_SPMV = 0,      // y = A x (r -> x), optional p'Ap
    BK_FWD_ZERO,
    BK_BWD_ZERO,
    BK_FWD_FULL,
    BK_BWD_FULL,
    BK_RESTRICT,
    BK_PROLONG,
    BK_DOT,           // reduction of a . b
    BK_CG_P,          // p = z (+ beta p)
    BK_CG_XR,         // x += alpha p, r -= alpha Ap, ||r||
    BK_COPY,          // w = a
    BK_WAXPBY,        // w = alpha a + beta b (alpha == 1 form only: b - A x)
};

struct BLaunch
{
    int kind;
    int prec;         // PREC_F64 / PREC_MIXED (value and vector types)
    int r_f64;        // rhs is f64 (mixed level 0)
    int dot;          // SpMV: fuse p'Ap
    int bf32;         // vector kernels: the second operand (z) is f32
    int block;        // threads per block
    int win;          // sweeps: widest slot window of the launch's segments (Level::win)
    int win_tail;     // post-smoother forward sweep: widest last-colour window
    int all_u;        // post-smoother forward sweep: every segment reuses its upper sums
    int rows;         // sweeps: rows per thread (1 or 2)
    int nseg;
    int64_t nblocks;
    BSeg* segs;       // device table
    double alpha, beta;
    int64_t nrows;    // rows (elements) over all segments
    int any_t;        // some segment keeps or reads SymGS partial sums (BSeg::t)
};

constexpr int BK_COUNT = 12;

struct Batch
{
    std::vector<Problem*> probs;
    int nlevels = 0;
    hipStream_t s = nullptr;
    bool own_stream = false;
    int block[MAX_LEVELS] = {256, 256, 256, 256, 256, 256, 256, 256};
    int rows[MAX_LEVELS] = {1, 1, 1, 1, 1, 1, 1, 1};   // sweeps: rows per thread, per level
    std::vector<BSeg*> tables;              // device allocations of the plans
    std::vector<BLaunch> pre[2], iter[2];   // preamble and one iteration, per CG mode
    bool planned = false;
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    int host_iter = 0;
    int mode = 0;
};

int batch_create(Batch*& B, int nprob, const int* dims, int nlevels, bool with_f64, bool with_mixed, hipStream_t s);
void batch_destroy(Batch* B);
int batch_set_block(Batch& B, int level, int block);
// rows per thread (1 or 2) of the sweep launches of a level; a launch whose
// segments are not all even in offset and size runs with 1
int batch_set_rows(Batch& B, int level, int rows);
int batch_cg(Batch& B, int mode, int max_iter, double tol, bool use_graph, int* niters, double* normr,
             double* normr0);
// `iters` batched CG iterations from each problem's x, every launch between
// HIP events: kind_ms[k] = summed milliseconds of the launches of kind k
// (BKind) over the iterations, kind_bytes[k] = their algorithmic HBM bytes per
// iteration (each slot a kernel loads once, gathers from cache, own-row
// vectors once), *iter_ms = milliseconds per iteration (first event to last)
int batch_profile(Batch& B, int mode, int iters, double* kind_ms, double* kind_bytes, double* iter_ms);

}  // namespace hpcgx
