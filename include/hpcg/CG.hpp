// CG.hpp -- preconditioned CG (CG.hpp:50-52, CG.cpp:104-200).  The loop keeps
// alpha/beta/norms on the device.  With times == NULL each iteration is
// replayed as a hipGraph; otherwise the kernels are launched one by one with
// HIP events around each phase, and times[] receives what the reference's
// timers hold (CG.cpp:187-198): [0] wall seconds, [1] dot products, [2]
// vector updates, [3] SpMV, [4] all-reduce (0 on one rank), [5] MG.  The
// p'Ap and ||r|| reductions are fused into the SpMV and the r update and are
// counted there.
#ifndef HPCGX_CG_HPP
#define HPCGX_CG_HPP
#include "CGData.hpp"
#include "SparseMatrix.hpp"
#include "Vector.hpp"
int CG(const SparseMatrix& A, CGData& data, const Vector& b, Vector& x, const int max_iter, const double tolerance,
       int& niters, double& normr, double& normr0, double* times, bool doPreconditioning, bool verbose);
#endif
