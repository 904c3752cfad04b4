// ComputeSPMV.hpp -- y = A x (ComputeSPMV.hpp:20, ComputeSPMV.cpp:226-278).
// When &y == A.mgData->Axf only the rows injected into the coarse level are
// computed (ComputeSPMV.cpp:238-256).
#ifndef HPCGX_COMPUTESPMV_HPP
#define HPCGX_COMPUTESPMV_HPP
#include "SparseMatrix.hpp"
int ComputeSPMV(const SparseMatrix& A, Vector& x, Vector& y);
#endif
