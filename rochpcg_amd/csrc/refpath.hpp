// refpath.hpp -- the reference's natural-order path (CG_ref and its _ref
// operators) on the device, for the benchmark driver's reference tolerance
// and as a GPU form of the reference's CPU configuration.
#pragma once
#include "solver.hpp"

namespace hpcgx
{

// Natural-order work vectors of a problem, allocated at the first use of the
// reference path and freed with the problem.
struct RefPath
{
    double *b = nullptr, *x = nullptr, *r = nullptr, *z = nullptr, *p = nullptr, *Ap = nullptr;
    double* lr[MAX_LEVELS] = {};   // levels >= 1: MGData::rc (natural order)
    double* lx[MAX_LEVELS] = {};   // levels >= 1: MGData::xc
    CGScalars* sc = nullptr;
    double* hist = nullptr;
    int hist_cap = 0;
};

void refpath_free(Problem& P);
// y = A x on level l (ComputeSPMV_ref.cpp:47-72)
int ref_spmv(Problem& P, int l, const double* x, double* y);
// one forward + one backward Gauss-Seidel sweep in natural row order (ComputeSYMGS_ref.cpp:54-103)
int ref_symgs(Problem& P, int l, const double* r, double* x);
// V-cycle from level l (ComputeMG_ref.cpp:39-63)
int ref_mg(Problem& P, int l, const double* r, double* x);
// CG_ref (CG_ref.cpp:86-174); b NULL = the problem's b taken to natural order,
// x NULL = the path's own x, zeroed first
int ref_cg(Problem& P, const double* b, double* x, int max_iter, double tol, bool precond, int* niters,
           double* normr, double* normr0, double* history, int hist_len);

}  // namespace hpcgx
