// ComputeMG.hpp -- V-cycle from A's level (ComputeMG.hpp:20, ComputeMG.cpp:64-102).
#ifndef HPCGX_COMPUTEMG_HPP
#define HPCGX_COMPUTEMG_HPP
#include "SparseMatrix.hpp"
int ComputeMG(const SparseMatrix& A, const Vector& r, Vector& x);
#endif
