// ComputeSYMGS.hpp -- one multicolour symmetric Gauss-Seidel step
// (ComputeSYMGS.hpp:49-50, ComputeSYMGS.cpp:369-454).
#ifndef HPCGX_COMPUTESYMGS_HPP
#define HPCGX_COMPUTESYMGS_HPP
#include "SparseMatrix.hpp"
int ComputeSYMGS(const SparseMatrix& A, const Vector& r, Vector& x);
int ComputeSYMGSZeroGuess(const SparseMatrix& A, const Vector& r, Vector& x);
#endif
