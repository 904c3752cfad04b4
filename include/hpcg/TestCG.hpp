// TestCG.hpp -- the result record and entry point of the reference's TestCG
// (TestCG.hpp:29-39 of ROCm/rocHPCG).  The layer does not implement TestCG:
// the reference's TestCG.cpp compiles against these headers and runs on the
// layer's CG, HIPCopyMatrixDiagonal, HIPReplaceMatrixDiagonal and vector
// helpers, reading SparseMatrix::d_localToGlobalMap.
#ifndef HPCGX_TESTCG_HPP
#define HPCGX_TESTCG_HPP

#include "CGData.hpp"
#include "SparseMatrix.hpp"
#include "Vector.hpp"
#include "hpcg.hpp"

struct TestCGData_STRUCT
{
    int count_pass;
    int count_fail;
    int expected_niters_no_prec;
    int expected_niters_prec;
    int niters_max_no_prec;
    int niters_max_prec;
    double normr;
};
typedef struct TestCGData_STRUCT TestCGData;

extern int TestCG(SparseMatrix& A, CGData& data, Vector& b, Vector& x, TestCGData& testcg_data);

#endif
