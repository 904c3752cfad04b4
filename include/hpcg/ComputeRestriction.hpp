// ComputeRestriction.hpp -- coarse residual at the injected points
// (ComputeRestriction.hpp:35-36): the Axf form and the fused SpMV form.
#ifndef HPCGX_COMPUTERESTRICTION_HPP
#define HPCGX_COMPUTERESTRICTION_HPP
#include "SparseMatrix.hpp"
int ComputeRestriction(const SparseMatrix& A, const Vector& rf);
int ComputeFusedSpMVRestriction(const SparseMatrix& A, const Vector& rf, Vector& xf);
#endif
