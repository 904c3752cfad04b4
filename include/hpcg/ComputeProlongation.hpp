// ComputeProlongation.hpp -- xf += injection of xc (ComputeProlongation.hpp:35).
#ifndef HPCGX_COMPUTEPROLONGATION_HPP
#define HPCGX_COMPUTEPROLONGATION_HPP
#include "SparseMatrix.hpp"
int ComputeProlongation(const SparseMatrix& Af, Vector& xf);
#endif
