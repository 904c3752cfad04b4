// TestSymmetry.hpp -- the result record and entry point of the reference's
// TestSymmetry (TestSymmetry.hpp:27-34 of ROCm/rocHPCG).  The layer does not
// implement TestSymmetry: the reference's TestSymmetry.cpp compiles against
// these headers and runs on the layer's ComputeSPMV, ComputeMG,
// ComputeDotProduct and ComputeResidual.
#ifndef HPCGX_TESTSYMMETRY_HPP
#define HPCGX_TESTSYMMETRY_HPP

#include "CGData.hpp"
#include "SparseMatrix.hpp"
#include "hpcg.hpp"

struct TestSymmetryData_STRUCT
{
    double depsym_spmv;
    double depsym_mg;
    int count_fail;
};
typedef struct TestSymmetryData_STRUCT TestSymmetryData;

extern int TestSymmetry(SparseMatrix& A, Vector& b, Vector& xexact, TestSymmetryData& testsymmetry_data);

#endif
