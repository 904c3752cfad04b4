// ComputeWAXPBY.hpp -- w = alpha x + beta y, and the fused y += alpha x with
// ||y||^2 (ComputeWAXPBY.hpp:48-61).
#ifndef HPCGX_COMPUTEWAXPBY_HPP
#define HPCGX_COMPUTEWAXPBY_HPP
#include "Vector.hpp"
int ComputeWAXPBY(const local_int_t n, const double alpha, const Vector& x, const double beta, const Vector& y,
                  Vector& w, bool& isOptimized);
int ComputeFusedWAXPBYDot(const local_int_t n, const double alpha, const Vector& x, Vector& y, double& result,
                          double& time_allreduce);
#endif
