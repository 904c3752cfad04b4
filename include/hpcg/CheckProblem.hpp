// CheckProblem.hpp -- CheckProblem.cpp:78 of the reference (called on every
// level by main.cpp:236-246).  Here it checks A's level in its optimized form
// on the device (every row: the in-range 27-point neighbours through the
// permutation, in order, 26 / -1, the closed-form counts, plus the colouring
// and layout; b, x, xexact when all three are given), so it runs after
// OptimizeProblem.  The reference asserts; this prints the mismatches to
// stderr.  HPCGX_CheckProblem (Extensions.hpp) returns them.
#ifndef HPCGX_CHECKPROBLEM_HPP
#define HPCGX_CHECKPROBLEM_HPP

#include "SparseMatrix.hpp"
#include "Vector.hpp"

void CheckProblem(SparseMatrix& A, Vector* b, Vector* x, Vector* xexact);

#endif
