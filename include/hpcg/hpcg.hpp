// hpcg.hpp -- umbrella header of the drop-in C++ layer.
#ifndef HPCGX_HPCG_HPP
#define HPCGX_HPCG_HPP
#include "CG.hpp"
#include "CGData.hpp"
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
#endif
