// ComputeDotProduct.hpp -- result = x'y over n entries (ComputeDotProduct.hpp:18-19).
// Synchronous (the result lands on the host); isOptimized is left unchanged.
#ifndef HPCGX_COMPUTEDOTPRODUCT_HPP
#define HPCGX_COMPUTEDOTPRODUCT_HPP
#include "Vector.hpp"
int ComputeDotProduct(const local_int_t n, const Vector& x, const Vector& y, double& result, double& time_allreduce,
                      bool& isOptimized);
#endif
