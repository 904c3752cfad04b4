// GenerateProblem.hpp -- GenerateProblem.cpp:375 / SetupHalo.cpp:313 /
// GenerateCoarseProblem.cpp:98 / OptimizeProblem.cpp:69 of the reference,
// single rank, built on the device.
//   GenerateProblem: geometry-derived sizes of A; b = 27 - nnz_row, x = 0,
//     xexact = 1 as device vectors in natural order.
//   SetupHalo: nothing to do on one rank.
//   GenerateCoarseProblem: the next coarser level (dims / 2) with its MGData.
//   OptimizeProblem: builds every level's multicolour ELL operator (device
//     JPL colouring, colour-stable permutation), permutes b, x, xexact into
//     multicolour order (replacing their d_values, as Permute.cpp:237-246
//     does) and allocates the CG vectors.  Returns non-zero on failure.
#ifndef HPCGX_GENERATEPROBLEM_HPP
#define HPCGX_GENERATEPROBLEM_HPP

#include "CGData.hpp"
#include "SparseMatrix.hpp"
#include "Vector.hpp"

void GenerateProblem(SparseMatrix& A, Vector* b, Vector* x, Vector* xexact);
void SetupHalo(SparseMatrix& A);
void GenerateCoarseProblem(const SparseMatrix& A);
int OptimizeProblem(SparseMatrix& A, CGData& data, Vector& b, Vector& x, Vector& xexact);

#endif
