// setup.hpp -- device-side construction of the multicolour hierarchy.
#pragma once
#include "common.hpp"

namespace hpcgx
{
// JPL colouring + colour-stable permutation of one level (fills colors, perm,
// ncolors/sizes/offsets/round_sizes, and a temporary inverse permutation).
int color_level(Level& L, hipStream_t s);
// Colouring plus the multicolour column-major ELL operator of an nx*ny*nz grid.
int build_level(Level& L, int nx, int ny, int nz, bool with_f64, bool with_f16, hipStream_t s);
// Injection maps from fine level F to the next coarser level C.
int build_transfer(Level& F, Level& C, hipStream_t s);
// b / xexact / x0 in multicolour order (kind 0: xexact = 1; kind 1: seeded uniform [-1,1]).
int make_rhs(const Level& L, int kind, uint64_t seed, double* b, double* xexact, double* x0, hipStream_t s);
void free_level(Level& L);
// natural-order helpers of the drop-in C++ layer
int natural_transfer(int nxf, int nyf, int nzf, int32_t* f2c, int32_t* c2f, hipStream_t s);
int permute_vector(int64_t m, const int32_t* perm, const double* in, double* out, hipStream_t s);
// p[i] = i for i < n
int iota_i64(int64_t n, int64_t* p, hipStream_t s);
int natural_rhs(int nx, int ny, int nz, double* b, double* xexact, double* x0, hipStream_t s);
}  // namespace hpcgx
