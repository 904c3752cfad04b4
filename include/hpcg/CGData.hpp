// CGData.hpp -- CG work vectors (CGData.hpp:55-107 of the reference).
#ifndef HPCGX_CGDATA_HPP
#define HPCGX_CGDATA_HPP

#include "SparseMatrix.hpp"
#include "Vector.hpp"

struct CGData_STRUCT
{
    Vector r;
    Vector z;
    Vector p;
    Vector Ap;
};
typedef struct CGData_STRUCT CGData;

void HIPInitializeSparseCGData(SparseMatrix& A, CGData& data);
void HIPDeleteCGData(CGData& data);

#endif
