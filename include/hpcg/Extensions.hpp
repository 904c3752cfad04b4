// Extensions.hpp -- functions of the drop-in layer with no counterpart in the
// reference.  The benchmark driver (rochpcg_amd/app/xhpcg.cpp) uses them for
// what the reference does with its own GPU kernels or host code.
#ifndef HPCGX_EXTENSIONS_HPP
#define HPCGX_EXTENSIONS_HPP

#include "CGData.hpp"
#include "SparseMatrix.hpp"
#include "Vector.hpp"

// Selects GPU `device` and creates the layer's stream and reduction scratch,
// what the reference does in HPCG_Init before its setup timer starts
// (init.cpp:193-233: device check, hipSetDevice, stream creation, workspace).
// Returns non-zero for a device id the machine does not have.  Without this
// call the layer starts on the current device at its first use.
int HPCGX_Init(int device);

// CG_ref (CG_ref.cpp:86-174) on the device, for the reference tolerance of
// main.cpp:334-344: natural row order, b = the problem's right-hand side
// (GenerateProblem's A*ones), x = 0, the reference's unoptimized operators
// with its rounding (ComputeSPMV_ref, ComputeSYMGS_ref as sequential
// Gauss-Seidel, ComputeMG_ref); only the dot products are summed in another
// order.  times[0] += wall seconds.
int HPCGX_ReferenceCG(const SparseMatrix& A, int max_iter, double tolerance, int& niters, double& normr,
                      double& normr0, double* times);

// CheckProblem's statements (CheckProblem.cpp:78-179) checked on the device
// over every row of A's level in its optimized form: the permutation, the
// colouring, each ELL row's entries (exactly the in-range 27-point neighbours
// in ascending order, 26 / -1, the closed-form count), the diagonal arrays and
// row descriptors, the level's nonzero total, and -- when b, x and xexact are
// all given -- b = 26 - (nnz_row - 1), x = 0, xexact = 1.  Returns the number
// of mismatches (0: the problem is as generated), or -1 when A is not
// optimized or a vector does not match it.  counts (may be null) receives 8
// entries: mismatches in the permutation, colouring, ELL rows, values,
// descriptors and vectors, then the stored nonzeros and the rows checked.
long long HPCGX_CheckProblem(const SparseMatrix& A, const Vector* b, const Vector* x, const Vector* xexact,
                             long long* counts);

// TestCG's exaggeration of the diagonal and of b (TestCG.cpp:67-95): the entry
// at position i is scaled by (g+2)*1e6 when its global row id g < 9 and by 1e6
// otherwise.  Like the reference, g is the natural row id of position i
// (d_localToGlobalMap is not permuted), so on one rank positions 0..8 take the
// large factors.
int HPCGX_TestCGScale(const SparseMatrix& A, Vector& diagonal, Vector& b);

// Runs `iterations` preconditioned FP64 CG iterations on (A, b) from x, one
// launch at a time with HIP events around each phase, and adds the seconds
// of each phase into times[] in the layout of the reference's CG timers
// (CG.cpp:68-82): times[0] total, [1] dot products, [2] vector updates,
// [3] SpMV, [5] MG.
int HPCGX_CGPhaseTimes(const SparseMatrix& A, CGData& data, const Vector& b, Vector& x, int iterations,
                       double* times);

// Device bytes held by the solver of A's hierarchy (matrices, transfer maps,
// vectors and scratch), after OptimizeProblem.
double HPCGX_DeviceBytes(const SparseMatrix& A);

// Name of the GPU the library runs on (a short string in buf).
int HPCGX_DeviceName(char* buf, int len);

#endif
