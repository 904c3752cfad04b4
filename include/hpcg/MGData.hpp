// MGData.hpp -- multigrid transfer data of the drop-in C++ layer
// (MGData.hpp:59-120 of the reference).
#ifndef HPCGX_MGDATA_HPP
#define HPCGX_MGDATA_HPP

#include "Vector.hpp"

struct MGData_STRUCT
{
    int numberOfPresmootherSteps;
    int numberOfPostsmootherSteps;
    local_int_t* f2cOperator;      // host copy (not kept by this layer)
    Vector* rc;                    // coarse residual (device buffer owned by the solver after OptimizeProblem)
    Vector* xc;                    // coarse solution (idem)
    Vector* Axf;                   // fine A*x at the injected rows (ComputeSPMV on Axf, ComputeRestriction)
    void* optimizationData;
    local_int_t* d_f2cOperator;    // coarse natural row -> fine natural row
    local_int_t* d_c2fOperator;    // fine natural row -> coarse natural row, -1 elsewhere
    bool owns_vectors;             // false once the solver owns rc/xc
};
typedef struct MGData_STRUCT MGData;

void InitializeMGData(local_int_t* d_f2cOperator, local_int_t* d_c2fOperator, Vector* rc, Vector* xc, Vector* Axf,
                      MGData& data);
void DeleteMGData(MGData& data);

#endif
