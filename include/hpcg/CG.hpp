// CG.hpp -- preconditioned CG (CG.hpp:50-52, CG.cpp:104-200).  The loop keeps
// alpha/beta/norms on the device and replays each iteration as a hipGraph;
// times[0] receives the wall time, times[1..5] the event-timed phase sums when
// the solver's phase profiling is requested through hpcgx_cg_iterate_profiled.
#ifndef HPCGX_CG_HPP
#define HPCGX_CG_HPP
#include "CGData.hpp"
#include "SparseMatrix.hpp"
#include "Vector.hpp"
int CG(const SparseMatrix& A, CGData& data, const Vector& b, Vector& x, const int max_iter, const double tolerance,
       int& niters, double& normr, double& normr0, double* times, bool doPreconditioning, bool verbose);
#endif
