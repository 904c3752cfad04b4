/*
 * hpcg_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the HPCG reference ("_ref") path of ROCm/rocHPCG, used as
 * the checker for the HIP library and as the CPU baseline. Nothing in the
 * product (rochpcg_amd/) links, loads or calls this code.
 *
 * Each function cites the reference file:line whose behaviour it restates.
 * Parity pins for every function are listed at the top of hpcg_oracle.c.
 */
#ifndef HPCG_ORACLE_H
#define HPCG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_W 27          /* numberOfNonzerosPerRow, GenerateProblem_ref.cpp:73 */
#define ORC_MAXLEV 8
#define ORC_MAXCOLORS 128 /* MAX_COLORS, utils.hpp:46 */

typedef struct orc_problem orc_problem;

/* threads used by every OpenMP loop of the oracle (0 = OpenMP default) */
void orc_set_threads(int nthreads);
int orc_get_threads(void);

/* --- problem construction ------------------------------------------------ */
/* order: 0 = natural row order (the _ref path), 1 = multicolor order built by
 * the JPL simulation, 2 = multicolor order from the closed-form colour table. */
orc_problem* orc_create(int nx, int ny, int nz, int nlevels, int order);
void orc_destroy(orc_problem* p);
int orc_nlevels(const orc_problem* p);
/* info[0..6] = nx, ny, nz, m, mc (coarse rows, 0 at coarsest), ncolors, nnz_total_lo32
 * returns the 64-bit nonzero count */
int64_t orc_level_info(const orc_problem* p, int level, int* info);
/* copies of the level's arrays (row-major [row*27+k]); any pointer may be NULL */
int orc_export_level(const orc_problem* p, int level, int* cols, double* vals,
                     unsigned char* nnz, int* diag, int* perm, int* f2c,
                     int* sizes, int* offsets, int* colors);

/* --- pieces of the setup, exposed for closed-form checks ------------------- */
void orc_generate(int nx, int ny, int nz, int* cols, double* vals,
                  unsigned char* nnz, int* diag, double* b, double* xexact);
void orc_f2c(int nxf, int nyf, int nzf, int* f2c);
void orc_row_hash(int nx, int ny, int nz, int* hash);
/* JPL colouring simulation over a row-major 27-wide adjacency; returns #colours */
int orc_jpl(int m, const int* cols, const int* hash, unsigned int seed, int* colors,
            int* round_sizes);
int orc_closed_form_colors(int nx, int ny, int nz, int* colors);
int orc_perm_from_colors(int m, const int* colors, int ncolors, int* perm,
                         int* sizes, int* offsets);
int orc_permute_matrix(int m, const int* cols, const double* vals,
                       const unsigned char* nnz, const int* perm, int* pcols,
                       double* pvals, unsigned char* pnnz, int* pdiag);

/* --- vectors in the problem's row order ----------------------------------- */
/* b = 27 - nnz_row (natural rows), permuted into the problem's order */
void orc_rhs_ones(const orc_problem* p, double* b);
/* x*_nat[i] = 2u-1, u = splitmix64 output i of `seed`; b_nat = A x* summed in
 * natural stencil order; both returned in the problem's order */
void orc_rhs_random(const orc_problem* p, uint64_t seed, double* xstar, double* b);
double orc_uniform(uint64_t seed, int64_t i);
void orc_permute_vector(const orc_problem* p, int level, const double* in_nat, double* out);
void orc_unpermute_vector(const orc_problem* p, int level, const double* in, double* out_nat);

/* --- the _ref kernels ----------------------------------------------------- */
int orc_spmv(const orc_problem* p, int level, const double* x, double* y);
int orc_symgs(const orc_problem* p, int level, const double* r, double* x);
int orc_mg(orc_problem* p, int level, const double* r, double* x);
double orc_dot(int n, const double* x, const double* y);
/* 0 = the reference's OpenMP reduction (default), 1 = compensated summation */
void orc_set_dot_mode(int mode);
int orc_waxpby(int n, double alpha, const double* x, double beta, const double* y, double* w);
int orc_restriction(orc_problem* p, int level, const double* rf);
int orc_prolongation(orc_problem* p, int level, double* xf);
/* Axf / rc / xc scratch of a level, for inspection */
int orc_get_level_vectors(const orc_problem* p, int level, double* Axf, double* rc, double* xc);

int orc_cg(orc_problem* p, const double* b, double* x, int max_iter, double tolerance,
           int precond, int* niters, double* normr, double* normr0, double* history,
           double* times);

/* diagonal access (CopyMatrixDiagonal / ReplaceMatrixDiagonal, SparseMatrix.hpp:233-255) */
int orc_copy_diagonal(const orc_problem* p, int level, double* d);
int orc_replace_diagonal(orc_problem* p, int level, const double* d);

#ifdef __cplusplus
}
#endif
#endif
