// hpcg.hpp -- umbrella header of the drop-in C++ layer.
#ifndef HPCGX_HPCG_HPP
#define HPCGX_HPCG_HPP
#include <fstream>
#include <iostream>
#include "CG.hpp"
#include "CGData.hpp"
#include "CheckProblem.hpp"
#include "ComputeDotProduct.hpp"
#include "ComputeMG.hpp"
#include "ComputeProlongation.hpp"
#include "ComputeRestriction.hpp"
#include "ComputeSPMV.hpp"
#include "ComputeSYMGS.hpp"
#include "ComputeWAXPBY.hpp"
#include "GenerateProblem.hpp"
#include "Geometry.hpp"
#include "MGData.hpp"
#include "SparseMatrix.hpp"
#include "Vector.hpp"
#include "ComputeResidual.hpp"

// the run log the reference's V&V code writes to (hpcg.hpp:55 declares it,
// init.cpp opens it); this layer defines it unopened -- a driver that wants
// the log opens it (rochpcg_amd/app/xhpcg.cpp writes its own)
extern std::ofstream HPCG_fout;
using std::endl;
#endif
