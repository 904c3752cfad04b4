// ComputeResidual.hpp -- max_i |v1_i - v2_i| over the first n entries of two
// device vectors (ComputeResidual.hpp:18, ComputeResidual.cpp:121-152): the
// inf-norm residual TestSymmetry reports for A*xexact against b.  Returns
// after the stream drains, with the value in `residual`.
#ifndef HPCGX_COMPUTERESIDUAL_HPP
#define HPCGX_COMPUTERESIDUAL_HPP
#include "Vector.hpp"
int ComputeResidual(const local_int_t n, const Vector& v1, const Vector& v2, double& residual);
#endif
