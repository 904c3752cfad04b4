/*
 * hpcg_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * A clean-room CPU restatement of the HPCG reference path of ROCm/rocHPCG
 * (the "_ref" kernels plus the setup that feeds them).  It is the checker for
 * the HIP library (tests/, __graft_entry__.smoke) and the CPU baseline of
 * bench.py.  The product never links or calls it.
 *
 * Numerics: built with -ffp-contract=off (no FMA contraction, no fast-math) so
 * that every expression is evaluated exactly as written, in the order the
 * reference's loops write it.  OpenMP loops mirror the reference's pragmas; the
 * dot-product reduction therefore depends on the thread count exactly as the
 * reference's does (run with 1 thread for reproducible fixtures).
 *
 * Parity pins (tests/test_oracle.py, tests/test_oracle_ref.py):
 *   orc_generate         closed forms: nnz_row = prod(1+[i>0]+[i<n-1]),
 *                        nnz_total = prod(3n-2), b = 27-nnz_row, A*ones = b,
 *                        symmetry, brute-force dense stencil at 2^3..4^3;
 *                        bit-exact vs the reference's GenerateProblem_ref
 *                        built in oracle/_ref.
 *   orc_f2c              closed form 2*ic per dimension; bijection onto the
 *                        all-even points.
 *   orc_row_hash         among neighbours, hash order == parity-class order
 *                        (where the grid keeps neighbouring hashes apart);
 *                        the JPL colouring it drives equals the closed-form
 *                        class->colour table below.
 *   orc_jpl              closed-form colour table (class->colour
 *                        {7:0,6:1,1:2,2:3,5:4,4:5,3:6,0:7}) derived from the
 *                        glibc rand() draws 0,7,1,2,4,3,5,6; proper colouring
 *                        (no neighbour shares a colour) on uneven grids.
 *   orc_perm_from_colors stable counting sort == numpy stable argsort.
 *   orc_permute_matrix   P A P^T == dense permutation of the dense matrix.
 *   orc_spmv/symgs/mg/   bit-exact vs the reference's own _ref sources built
 *   dot/waxpby/cg        in oracle/_ref (same compiler flags); plus dense
 *                        brute force (SpMV), the triangular-solve identity
 *                        (SymGS), x'M^-1y symmetry (MG).
 */
#include "hpcg_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    int nx, ny, nz;
    int m;
    int64_t nnz_total;
    int* cols;          /* [m*27] row-major, -1 padding at the tail */
    double* vals;       /* [m*27], 0.0 padding */
    unsigned char* nnz; /* nonzerosInRow */
    int* diag;          /* slot of the diagonal entry (matrixDiagonal) */
    double* Axf;        /* MGData::Axf */
    int mc;             /* coarse rows */
    int* f2c;           /* coarse row -> fine row, both in this problem's order */
    double* rc;         /* MGData::rc */
    double* xc;         /* MGData::xc */
    int* perm;          /* natural row -> row in this problem's order (NULL: natural) */
    int* colors;        /* colour of each natural row (NULL: natural) */
    int ncolors;
    int sizes[ORC_MAXCOLORS];
    int offsets[ORC_MAXCOLORS + 1];
} orc_level;

struct orc_problem {
    int nlevels;
    int order;
    orc_level L[ORC_MAXLEV];
};

static int g_threads = 0;

void orc_set_threads(int nthreads)
{
    g_threads = nthreads;
    if(nthreads > 0) omp_set_num_threads(nthreads);
}

int orc_get_threads(void) { return omp_get_max_threads(); }

/* ------------------------------------------------------------------------ */
/* Problem generation: GenerateProblem_ref.cpp:131-182 (stencil order sz, sy,
 * sx in -1..1, in-bounds only; diagonal 26, others -1; b = 26-(nnz-1);
 * x = 0; xexact = 1).                                                       */
void orc_generate(int nx, int ny, int nz, int* cols, double* vals, unsigned char* nnz,
                  int* diag, double* b, double* xexact)
{
    const int64_t nxy = (int64_t)nx * ny;
#pragma omp parallel for schedule(static)
    for(int iz = 0; iz < nz; ++iz)
    {
        for(int iy = 0; iy < ny; ++iy)
        {
            for(int ix = 0; ix < nx; ++ix)
            {
                const int64_t row = iz * nxy + (int64_t)iy * nx + ix;
                int* ci = cols + row * ORC_W;
                double* vi = vals + row * ORC_W;
                int k = 0;
                for(int sz = -1; sz <= 1; ++sz)
                {
                    if(iz + sz < 0 || iz + sz >= nz) continue;
                    for(int sy = -1; sy <= 1; ++sy)
                    {
                        if(iy + sy < 0 || iy + sy >= ny) continue;
                        for(int sx = -1; sx <= 1; ++sx)
                        {
                            if(ix + sx < 0 || ix + sx >= nx) continue;
                            int64_t col = row + sz * nxy + (int64_t)sy * nx + sx;
                            if(col == row)
                            {
                                if(diag) diag[row] = k;
                                vi[k] = 26.0;
                            }
                            else
                            {
                                vi[k] = -1.0;
                            }
                            ci[k] = (int)col;
                            ++k;
                        }
                    }
                }
                nnz[row] = (unsigned char)k;
                for(int j = k; j < ORC_W; ++j)
                {
                    ci[j] = -1;
                    vi[j] = 0.0;
                }
                if(b) b[row] = 26.0 - ((double)(k - 1));
                if(xexact) xexact[row] = 1.0;
            }
        }
    }
}

/* Injection operator: GenerateCoarseProblem.cpp:57-86 (kernel_f2c_operator). */
void orc_f2c(int nxf, int nyf, int nzf, int* f2c)
{
    const int nxc = nxf / 2, nyc = nyf / 2, nzc = nzf / 2;
    for(int izc = 0; izc < nzc; ++izc)
        for(int iyc = 0; iyc < nyc; ++iyc)
            for(int ixc = 0; ixc < nxc; ++ixc)
            {
                int64_t c = ((int64_t)izc * nyc + iyc) * nxc + ixc;
                int64_t f = (int64_t)(2 * izc) * nxf * nyf + (int64_t)(2 * iyc) * nxf + 2 * ixc;
                f2c[c] = (int)f;
            }
}

/* Row hash: GenerateProblem.cpp:93-96 (get_hash) and :294-295
 * hash = class*nx*ny*nz + iz*nx*ny + 2*iy*nx + 4*ix,
 * class = ((ix&1)<<2)|((iy&1)<<1)|(iz&1).                                    */
void orc_row_hash(int nx, int ny, int nz, int* hash)
{
    const int64_t nxy = (int64_t)nx * ny;
#pragma omp parallel for schedule(static)
    for(int iz = 0; iz < nz; ++iz)
        for(int iy = 0; iy < ny; ++iy)
            for(int ix = 0; ix < nx; ++ix)
            {
                int64_t row = iz * nxy + (int64_t)iy * nx + ix;
                int cls = ((ix & 1) << 2) | ((iy & 1) << 1) | (iz & 1);
                int64_t crd = iz * nxy + (int64_t)iy * (nx << 1) + (ix << 2);
                hash[row] = (int)(cls * nxy * nz + crd);
            }
}

/* Jones-Plassmann-Luby colouring: MultiColoring.cpp:151-236 (kernel_jpl) and
 * :238-327 (JPLColoring round loop).  A round colours an uncoloured row with
 * color1 if its hash exceeds every neighbour that is uncoloured or coloured
 * color1/color2, with color2 if it is below all of them.  Rows coloured in the
 * same round read as "uncoloured or color1/color2", so each round depends only
 * on the state before it; the simulation evaluates it on that state.
 * Equal hashes of neighbouring rows (e.g. a 4x3x2 level) make the reference
 * loop forever; ties are broken by row index here and on the device.  The
 * first 8 colours come from srand(0x586744); rand()%8 (MultiColoring.cpp:248,
 * 284-285), later ones count upwards.  round_sizes[k] is what the reference
 * stores in A.sizes[k] (count of rows holding that round's colour value).    */
int orc_jpl(int m, const int* cols, const int* hash, unsigned int seed, int* colors,
            int* round_sizes)
{
    int* next = (int*)malloc(sizeof(int) * (size_t)m);
    for(int i = 0; i < m; ++i) colors[i] = -1;
    srand(seed);
    int nblocks = 0;
    int64_t colored = 0;
    while(colored != m && nblocks + 2 <= ORC_MAXCOLORS)
    {
        const int color1 = (nblocks < 8) ? rand() % 8 : nblocks;
        const int color2 = (nblocks < 8) ? rand() % 8 : nblocks + 1;
#pragma omp parallel for schedule(static)
        for(int row = 0; row < m; ++row)
        {
            next[row] = colors[row];
            if(colors[row] != -1) continue;
            int is_max = 1, is_min = 1;
            const int rh = hash[row];
            for(int k = 0; k < ORC_W; ++k)
            {
                int col = cols[(int64_t)row * ORC_W + k];
                if(col < 0 || col >= m || col == row) continue;
                int cn = colors[col];
                if(cn == -1 || cn == color1 || cn == color2)
                {
                    int ch = hash[col];
                    /* the reference tests ch >= rh / ch <= rh; with equal
                     * hashes it never colours either row and never ends.
                     * Breaking the tie by row index changes nothing in any
                     * run that ends (no equal-hash neighbours exist there). */
                    if(ch > rh || (ch == rh && col > row)) is_max = 0;
                    if(ch < rh || (ch == rh && col < row)) is_min = 0;
                }
            }
            if(is_max)
                next[row] = color1;
            else if(is_min)
                next[row] = color2;
        }
        int64_t c1 = 0, c2 = 0;
        for(int row = 0; row < m; ++row)
        {
            colors[row] = next[row];
            c1 += (next[row] == color1);
            c2 += (next[row] == color2);
        }
        if(round_sizes)
        {
            round_sizes[nblocks] = (int)c1;
            round_sizes[nblocks + 1] = (int)c2;
        }
        colored += c1 + c2;
        nblocks += 2;
    }
    free(next);
    int ncolors = 0;
    for(int i = 0; i < m; ++i)
        if(colors[i] + 1 > ncolors) ncolors = colors[i] + 1;
    return (colored == m) ? ncolors : -1;
}

/* Closed form of the JPL result on a 27-point grid with all dimensions >= 2
 * (every row then has neighbours of all eight parity classes, so round r
 * colours class 7-r as maximum and class r as minimum).                      */
int orc_closed_form_colors(int nx, int ny, int nz, int* colors)
{
    static const int class_to_color[8] = {7, 2, 3, 6, 5, 4, 1, 0};
    const int64_t nxy = (int64_t)nx * ny;
    for(int iz = 0; iz < nz; ++iz)
        for(int iy = 0; iy < ny; ++iy)
            for(int ix = 0; ix < nx; ++ix)
            {
                int cls = ((ix & 1) << 2) | ((iy & 1) << 1) | (iz & 1);
                colors[iz * nxy + (int64_t)iy * nx + ix] = class_to_color[cls];
            }
    return 8;
}

/* Stable sort of rows by colour: MultiColoring.cpp:328-344 (rocPRIM
 * radix_sort_pairs over (colour, identity) is stable) and kernel_create_perm
 * (perm[sorted_old[i]] = i).  sizes/offsets are per colour value.            */
int orc_perm_from_colors(int m, const int* colors, int ncolors, int* perm, int* sizes,
                         int* offsets)
{
    int cnt[ORC_MAXCOLORS + 1];
    memset(cnt, 0, sizeof(cnt));
    for(int i = 0; i < m; ++i)
    {
        if(colors[i] < 0 || colors[i] >= ncolors) return -1;
        cnt[colors[i]]++;
    }
    int off[ORC_MAXCOLORS + 1];
    off[0] = 0;
    for(int c = 0; c < ncolors; ++c) off[c + 1] = off[c] + cnt[c];
    if(sizes)
        for(int c = 0; c < ncolors; ++c) sizes[c] = cnt[c];
    if(offsets)
        for(int c = 0; c <= ncolors; ++c) offsets[c] = off[c];
    int run[ORC_MAXCOLORS];
    for(int c = 0; c < ncolors; ++c) run[c] = off[c];
    for(int i = 0; i < m; ++i) perm[i] = run[colors[i]]++;
    return 0;
}

/* Row and column permutation: Permute.cpp:95-149 (kernel_perm_cols: columns
 * mapped through perm, each row's entries sorted ascending by new column,
 * invalid slots -> -1 / 0.0 at the tail), Permute.cpp:54-76 and :187-218
 * (rows scattered to perm[row]), SparseMatrix.cpp:361-389 (diagonal slot).   */
int orc_permute_matrix(int m, const int* cols, const double* vals, const unsigned char* nnz,
                       const int* perm, int* pcols, double* pvals, unsigned char* pnnz,
                       int* pdiag)
{
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(+ : bad)
    for(int i = 0; i < m; ++i)
    {
        const int64_t src = (int64_t)i * ORC_W;
        const int64_t dst = (int64_t)perm[i] * ORC_W;
        int key[ORC_W];
        double val[ORC_W];
        int n = 0;
        for(int k = 0; k < nnz[i]; ++k)
        {
            int col = cols[src + k];
            if(col < 0 || col >= m) continue;
            /* insertion into ascending order */
            int kk = perm[col];
            int j = n++;
            while(j > 0 && key[j - 1] > kk)
            {
                key[j] = key[j - 1];
                val[j] = val[j - 1];
                --j;
            }
            key[j] = kk;
            val[j] = vals[src + k];
        }
        int d = -1;
        for(int k = 0; k < ORC_W; ++k)
        {
            pcols[dst + k] = (k < n) ? key[k] : -1;
            pvals[dst + k] = (k < n) ? val[k] : 0.0;
            if(k < n && key[k] == perm[i]) d = k;
        }
        pnnz[perm[i]] = (unsigned char)n;
        if(pdiag) pdiag[perm[i]] = d;
        if(d < 0) bad++;
    }
    return bad ? -1 : 0;
}

/* ------------------------------------------------------------------------ */
static void* xmalloc(size_t n)
{
    void* p = malloc(n ? n : 1);
    return p;
}

static int build_level(orc_level* L, int nx, int ny, int nz, int order)
{
    memset(L, 0, sizeof(*L));
    L->nx = nx;
    L->ny = ny;
    L->nz = nz;
    const int64_t m64 = (int64_t)nx * ny * nz;
    if(m64 <= 0 || m64 > 0x7fffffff) return -1;
    const int m = (int)m64;
    L->m = m;
    L->nnz_total = (int64_t)(3 * nx - 2) * (3 * ny - 2) * (3 * nz - 2);
    int* cols = (int*)xmalloc(sizeof(int) * (size_t)m * ORC_W);
    double* vals = (double*)xmalloc(sizeof(double) * (size_t)m * ORC_W);
    unsigned char* nnz = (unsigned char*)xmalloc((size_t)m);
    int* diag = (int*)xmalloc(sizeof(int) * (size_t)m);
    if(!cols || !vals || !nnz || !diag) return -2;
    orc_generate(nx, ny, nz, cols, vals, nnz, diag, NULL, NULL);
    if(order == 0)
    {
        L->cols = cols;
        L->vals = vals;
        L->nnz = nnz;
        L->diag = diag;
        L->ncolors = 1;
        L->sizes[0] = m;
        L->offsets[0] = 0;
        L->offsets[1] = m;
    }
    else
    {
        L->colors = (int*)xmalloc(sizeof(int) * (size_t)m);
        L->perm = (int*)xmalloc(sizeof(int) * (size_t)m);
        if(order == 1)
        {
            int* hash = (int*)xmalloc(sizeof(int) * (size_t)m);
            orc_row_hash(nx, ny, nz, hash);
            L->ncolors = orc_jpl(m, cols, hash, 0x586744u, L->colors, NULL);
            free(hash);
        }
        else
        {
            L->ncolors = orc_closed_form_colors(nx, ny, nz, L->colors);
        }
        if(L->ncolors <= 0) return -3;
        orc_perm_from_colors(m, L->colors, L->ncolors, L->perm, L->sizes, L->offsets);
        L->cols = (int*)xmalloc(sizeof(int) * (size_t)m * ORC_W);
        L->vals = (double*)xmalloc(sizeof(double) * (size_t)m * ORC_W);
        L->nnz = (unsigned char*)xmalloc((size_t)m);
        L->diag = (int*)xmalloc(sizeof(int) * (size_t)m);
        int rc = orc_permute_matrix(m, cols, vals, nnz, L->perm, L->cols, L->vals, L->nnz,
                                    L->diag);
        free(cols);
        free(vals);
        free(nnz);
        free(diag);
        if(rc) return -4;
    }
    L->Axf = (double*)calloc((size_t)m, sizeof(double));
    return 0;
}

orc_problem* orc_create(int nx, int ny, int nz, int nlevels, int order)
{
    if(nlevels < 1 || nlevels > ORC_MAXLEV || nx < 1 || ny < 1 || nz < 1) return NULL;
    orc_problem* p = (orc_problem*)calloc(1, sizeof(orc_problem));
    p->nlevels = nlevels;
    p->order = order;
    int cx = nx, cy = ny, cz = nz;
    for(int l = 0; l < nlevels; ++l)
    {
        if(l > 0)
        {
            /* GenerateCoarseProblem.cpp:106-113: fine dims must be even */
            if(cx % 2 || cy % 2 || cz % 2)
            {
                orc_destroy(p);
                return NULL;
            }
            cx /= 2;
            cy /= 2;
            cz /= 2;
        }
        if(build_level(&p->L[l], cx, cy, cz, order) != 0)
        {
            p->nlevels = l + 1;
            orc_destroy(p);
            return NULL;
        }
    }
    /* transfer operators between consecutive levels, in each level's order */
    for(int l = 0; l + 1 < nlevels; ++l)
    {
        orc_level* F = &p->L[l];
        orc_level* C = &p->L[l + 1];
        F->mc = C->m;
        int* f2c_nat = (int*)xmalloc(sizeof(int) * (size_t)C->m);
        orc_f2c(F->nx, F->ny, F->nz, f2c_nat);
        F->f2c = (int*)xmalloc(sizeof(int) * (size_t)C->m);
        for(int k = 0; k < C->m; ++k)
        {
            int cf = C->perm ? C->perm[k] : k;
            int ff = F->perm ? F->perm[f2c_nat[k]] : f2c_nat[k];
            F->f2c[cf] = ff;
        }
        free(f2c_nat);
        F->rc = (double*)calloc((size_t)C->m, sizeof(double));
        F->xc = (double*)calloc((size_t)C->m, sizeof(double));
    }
    return p;
}

void orc_destroy(orc_problem* p)
{
    if(!p) return;
    for(int l = 0; l < p->nlevels; ++l)
    {
        orc_level* L = &p->L[l];
        free(L->cols);
        free(L->vals);
        free(L->nnz);
        free(L->diag);
        free(L->Axf);
        free(L->f2c);
        free(L->rc);
        free(L->xc);
        free(L->perm);
        free(L->colors);
    }
    free(p);
}

int orc_nlevels(const orc_problem* p) { return p->nlevels; }

int64_t orc_level_info(const orc_problem* p, int level, int* info)
{
    const orc_level* L = &p->L[level];
    if(info)
    {
        info[0] = L->nx;
        info[1] = L->ny;
        info[2] = L->nz;
        info[3] = L->m;
        info[4] = L->mc;
        info[5] = L->ncolors;
        info[6] = (int)(L->nnz_total & 0x7fffffff);
    }
    return L->nnz_total;
}

int orc_export_level(const orc_problem* p, int level, int* cols, double* vals,
                     unsigned char* nnz, int* diag, int* perm, int* f2c, int* sizes,
                     int* offsets, int* colors)
{
    const orc_level* L = &p->L[level];
    const size_t m = (size_t)L->m;
    if(cols) memcpy(cols, L->cols, sizeof(int) * m * ORC_W);
    if(vals) memcpy(vals, L->vals, sizeof(double) * m * ORC_W);
    if(nnz) memcpy(nnz, L->nnz, m);
    if(diag) memcpy(diag, L->diag, sizeof(int) * m);
    if(perm)
    {
        if(L->perm)
            memcpy(perm, L->perm, sizeof(int) * m);
        else
            for(size_t i = 0; i < m; ++i) perm[i] = (int)i;
    }
    if(f2c && L->f2c) memcpy(f2c, L->f2c, sizeof(int) * (size_t)L->mc);
    if(sizes) memcpy(sizes, L->sizes, sizeof(int) * (size_t)L->ncolors);
    if(offsets) memcpy(offsets, L->offsets, sizeof(int) * (size_t)(L->ncolors + 1));
    if(colors)
    {
        if(L->colors)
            memcpy(colors, L->colors, sizeof(int) * m);
        else
            memset(colors, 0, sizeof(int) * m);
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* splitmix64: output i of the generator seeded with `seed` is
 * mix(seed + (i+1)*0x9E3779B97F4A7C15); u = (z >> 11) * 2^-53.               */
double orc_uniform(uint64_t seed, int64_t i)
{
    uint64_t z = seed + (uint64_t)(i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

void orc_permute_vector(const orc_problem* p, int level, const double* in_nat, double* out)
{
    const orc_level* L = &p->L[level];
#pragma omp parallel for schedule(static)
    for(int i = 0; i < L->m; ++i) out[L->perm ? L->perm[i] : i] = in_nat[i];
}

void orc_unpermute_vector(const orc_problem* p, int level, const double* in, double* out_nat)
{
    const orc_level* L = &p->L[level];
#pragma omp parallel for schedule(static)
    for(int i = 0; i < L->m; ++i) out_nat[i] = in[L->perm ? L->perm[i] : i];
}

void orc_rhs_ones(const orc_problem* p, double* b)
{
    const orc_level* L = &p->L[0];
    const int nx = L->nx, ny = L->ny, nz = L->nz;
    double* bn = (double*)xmalloc(sizeof(double) * (size_t)L->m);
#pragma omp parallel for schedule(static)
    for(int iz = 0; iz < nz; ++iz)
        for(int iy = 0; iy < ny; ++iy)
            for(int ix = 0; ix < nx; ++ix)
            {
                int k = (1 + (ix > 0) + (ix < nx - 1)) * (1 + (iy > 0) + (iy < ny - 1)) *
                        (1 + (iz > 0) + (iz < nz - 1));
                bn[((int64_t)iz * ny + iy) * nx + ix] = 26.0 - ((double)(k - 1));
            }
    orc_permute_vector(p, 0, bn, b);
    free(bn);
}

void orc_rhs_random(const orc_problem* p, uint64_t seed, double* xstar, double* b)
{
    const orc_level* L = &p->L[0];
    const int nx = L->nx, ny = L->ny, nz = L->nz;
    const int64_t nxy = (int64_t)nx * ny;
    double* xn = (double*)xmalloc(sizeof(double) * (size_t)L->m);
    double* bn = (double*)xmalloc(sizeof(double) * (size_t)L->m);
#pragma omp parallel for schedule(static)
    for(int64_t i = 0; i < L->m; ++i) xn[i] = 2.0 * orc_uniform(seed, i) - 1.0;
    /* natural-order row sums in stencil order (sz, sy, sx ascending) */
#pragma omp parallel for schedule(static)
    for(int iz = 0; iz < nz; ++iz)
        for(int iy = 0; iy < ny; ++iy)
            for(int ix = 0; ix < nx; ++ix)
            {
                int64_t row = iz * nxy + (int64_t)iy * nx + ix;
                double sum = 0.0;
                for(int sz = -1; sz <= 1; ++sz)
                {
                    if(iz + sz < 0 || iz + sz >= nz) continue;
                    for(int sy = -1; sy <= 1; ++sy)
                    {
                        if(iy + sy < 0 || iy + sy >= ny) continue;
                        for(int sx = -1; sx <= 1; ++sx)
                        {
                            if(ix + sx < 0 || ix + sx >= nx) continue;
                            int64_t col = row + sz * nxy + (int64_t)sy * nx + sx;
                            double a = (col == row) ? 26.0 : -1.0;
                            sum = fma(a, xn[col], sum);
                        }
                    }
                }
                bn[row] = sum;
            }
    orc_permute_vector(p, 0, xn, xstar);
    orc_permute_vector(p, 0, bn, b);
    free(xn);
    free(bn);
}

/* ------------------------------------------------------------------------ */
/* ComputeSPMV_ref.cpp:47-72 */
int orc_spmv(const orc_problem* p, int level, const double* x, double* y)
{
    const orc_level* A = &p->L[level];
    const int nrow = A->m;
#pragma omp parallel for schedule(static)
    for(int i = 0; i < nrow; i++)
    {
        double sum = 0.0;
        const double* cur_vals = A->vals + (int64_t)i * ORC_W;
        const int* cur_inds = A->cols + (int64_t)i * ORC_W;
        const int cur_nnz = A->nnz[i];
        for(int j = 0; j < cur_nnz; j++) sum += cur_vals[j] * x[cur_inds[j]];
        y[i] = sum;
    }
    return 0;
}

/* ComputeSYMGS_ref.cpp:54-103: sequential forward sweep then backward sweep;
 * the diagonal is included in the j-loop and added back afterwards.          */
int orc_symgs(const orc_problem* p, int level, const double* r, double* x)
{
    const orc_level* A = &p->L[level];
    const int nrow = A->m;
    for(int i = 0; i < nrow; i++)
    {
        const double* cv = A->vals + (int64_t)i * ORC_W;
        const int* ci = A->cols + (int64_t)i * ORC_W;
        const int cn = A->nnz[i];
        const double d = cv[A->diag[i]];
        double sum = r[i];
        for(int j = 0; j < cn; j++) sum -= cv[j] * x[ci[j]];
        sum += x[i] * d;
        x[i] = sum / d;
    }
    for(int i = nrow - 1; i >= 0; i--)
    {
        const double* cv = A->vals + (int64_t)i * ORC_W;
        const int* ci = A->cols + (int64_t)i * ORC_W;
        const int cn = A->nnz[i];
        const double d = cv[A->diag[i]];
        double sum = r[i];
        for(int j = 0; j < cn; j++) sum -= cv[j] * x[ci[j]];
        sum += x[i] * d;
        x[i] = sum / d;
    }
    return 0;
}

/* ComputeRestriction_ref.cpp:40-54 */
int orc_restriction(orc_problem* p, int level, const double* rf)
{
    orc_level* A = &p->L[level];
    const int nc = A->mc;
#pragma omp parallel for schedule(static)
    for(int i = 0; i < nc; ++i) A->rc[i] = rf[A->f2c[i]] - A->Axf[A->f2c[i]];
    return 0;
}

/* ComputeProlongation_ref.cpp:38-52 */
int orc_prolongation(orc_problem* p, int level, double* xf)
{
    orc_level* A = &p->L[level];
    const int nc = A->mc;
#pragma omp parallel for schedule(static)
    for(int i = 0; i < nc; ++i) xf[A->f2c[i]] += A->xc[i];
    return 0;
}

/* ComputeMG_ref.cpp:39-63: zero x, 1 pre-smoothing SymGS, fine residual via a
 * full SpMV into Axf, injection, recursion, prolongation, 1 post-smoothing
 * SymGS; the coarsest level does one SymGS.                                  */
int orc_mg(orc_problem* p, int level, const double* r, double* x)
{
    orc_level* A = &p->L[level];
    memset(x, 0, sizeof(double) * (size_t)A->m);
    if(level + 1 < p->nlevels)
    {
        orc_symgs(p, level, r, x);
        orc_spmv(p, level, x, A->Axf);
        orc_restriction(p, level, r);
        orc_mg(p, level + 1, A->rc, A->xc);
        orc_prolongation(p, level, x);
        orc_symgs(p, level, r, x);
    }
    else
    {
        orc_symgs(p, level, r, x);
    }
    return 0;
}

int orc_get_level_vectors(const orc_problem* p, int level, double* Axf, double* rc, double* xc)
{
    const orc_level* A = &p->L[level];
    if(Axf) memcpy(Axf, A->Axf, sizeof(double) * (size_t)A->m);
    if(rc && A->rc) memcpy(rc, A->rc, sizeof(double) * (size_t)A->mc);
    if(xc && A->xc) memcpy(xc, A->xc, sizeof(double) * (size_t)A->mc);
    return 0;
}

/* Dot-product summation of the oracle: 0 = the reference's OpenMP reduction
 * (ComputeDotProduct_ref.cpp:46-80: each thread sums its block of the loop in
 * order, the partials are then added), 1 = compensated summation (Neumaier's
 * variant of Kahan summation, sequential) of the same rounded products, whose
 * error is about one rounding of the total.  Mode 1 is not the reference's
 * arithmetic: it measures how much of a history difference comes from the
 * reference's own summation error, which grows with the vector length.     */
static int g_dot_mode = 0;

void orc_set_dot_mode(int mode) { g_dot_mode = mode; }

static double dot_compensated(int n, const double* x, const double* y)
{
    double s = 0.0, c = 0.0;
    for(int i = 0; i < n; i++)
    {
        const double p = x[i] * y[i];
        const double t = s + p;
        if(fabs(s) >= fabs(p))
            c += (s - t) + p;
        else
            c += (p - t) + s;
        s = t;
    }
    return s + c;
}

/* ComputeDotProduct_ref.cpp:46-80 */
double orc_dot(int n, const double* x, const double* y)
{
    if(g_dot_mode == 1) return dot_compensated(n, x, y);
    double local_result = 0.0;
    if(y == x)
    {
#pragma omp parallel for reduction(+ : local_result)
        for(int i = 0; i < n; i++) local_result += x[i] * x[i];
    }
    else
    {
#pragma omp parallel for reduction(+ : local_result)
        for(int i = 0; i < n; i++) local_result += x[i] * y[i];
    }
    return local_result;
}

/* ComputeWAXPBY_ref.cpp:42-70 */
int orc_waxpby(int n, double alpha, const double* x, double beta, const double* y, double* w)
{
    if(alpha == 1.0)
    {
#pragma omp parallel for
        for(int i = 0; i < n; i++) w[i] = x[i] + beta * y[i];
    }
    else if(beta == 1.0)
    {
#pragma omp parallel for
        for(int i = 0; i < n; i++) w[i] = alpha * x[i] + y[i];
    }
    else
    {
#pragma omp parallel for
        for(int i = 0; i < n; i++) w[i] = alpha * x[i] + beta * y[i];
    }
    return 0;
}

/* CG_ref.cpp:86-174.  history[0] = normr0, history[k] = normr after
 * iteration k (the "Scaled Residual" log is history[k]/history[0]).
 * times[0..5] as in CG_ref (total, ddot, waxpby, spmv, allreduce, mg).       */
int orc_cg(orc_problem* p, const double* b, double* x, int max_iter, double tolerance,
           int precond, int* niters, double* normr_out, double* normr0_out, double* history,
           double* times)
{
    orc_level* A = &p->L[0];
    const int nrow = A->m;
    double t_begin = omp_get_wtime();
    double t0 = 0, t1 = 0, t2 = 0, t3 = 0, t5 = 0;
    double normr = 0.0, rtz = 0.0, oldrtz = 0.0, alpha = 0.0, beta = 0.0, pAp = 0.0;
    double* r = (double*)xmalloc(sizeof(double) * (size_t)nrow);
    double* z = (double*)xmalloc(sizeof(double) * (size_t)nrow);
    double* pv = (double*)xmalloc(sizeof(double) * (size_t)nrow);
    double* Ap = (double*)xmalloc(sizeof(double) * (size_t)nrow);
#define TICK() t0 = omp_get_wtime()
#define TOCK(t) t += omp_get_wtime() - t0
    memcpy(pv, x, sizeof(double) * (size_t)nrow);
    TICK();
    orc_spmv(p, 0, pv, Ap);
    TOCK(t3);
    TICK();
    orc_waxpby(nrow, 1.0, b, -1.0, Ap, r);
    TOCK(t2);
    TICK();
    normr = orc_dot(nrow, r, r);
    TOCK(t1);
    normr = sqrt(normr);
    double normr0 = normr;
    if(history) history[0] = normr;
    int k;
    *niters = 0;
    for(k = 1; k <= max_iter && normr / normr0 > tolerance; k++)
    {
        TICK();
        if(precond)
            orc_mg(p, 0, r, z);
        else
            orc_waxpby(nrow, 1.0, r, 0.0, r, z);
        TOCK(t5);
        if(k == 1)
        {
            memcpy(pv, z, sizeof(double) * (size_t)nrow);
            TOCK(t2);
            TICK();
            rtz = orc_dot(nrow, r, z);
            TOCK(t1);
        }
        else
        {
            oldrtz = rtz;
            TICK();
            rtz = orc_dot(nrow, r, z);
            TOCK(t1);
            beta = rtz / oldrtz;
            TICK();
            orc_waxpby(nrow, 1.0, z, beta, pv, pv);
            TOCK(t2);
        }
        TICK();
        orc_spmv(p, 0, pv, Ap);
        TOCK(t3);
        TICK();
        pAp = orc_dot(nrow, pv, Ap);
        TOCK(t1);
        alpha = rtz / pAp;
        TICK();
        orc_waxpby(nrow, 1.0, x, alpha, pv, x);
        orc_waxpby(nrow, 1.0, r, -alpha, Ap, r);
        TOCK(t2);
        TICK();
        normr = orc_dot(nrow, r, r);
        TOCK(t1);
        normr = sqrt(normr);
        if(history) history[k] = normr;
        *niters = k;
    }
#undef TICK
#undef TOCK
    if(times)
    {
        times[1] += t1;
        times[2] += t2;
        times[3] += t3;
        times[4] += 0.0;
        times[5] += t5;
        times[0] += omp_get_wtime() - t_begin;
    }
    *normr_out = normr;
    *normr0_out = normr0;
    free(r);
    free(z);
    free(pv);
    free(Ap);
    return 0;
}

int orc_copy_diagonal(const orc_problem* p, int level, double* d)
{
    const orc_level* A = &p->L[level];
    for(int i = 0; i < A->m; ++i) d[i] = A->vals[(int64_t)i * ORC_W + A->diag[i]];
    return 0;
}

int orc_replace_diagonal(orc_problem* p, int level, const double* d)
{
    orc_level* A = &p->L[level];
    for(int i = 0; i < A->m; ++i) A->vals[(int64_t)i * ORC_W + A->diag[i]] = d[i];
    return 0;
}
