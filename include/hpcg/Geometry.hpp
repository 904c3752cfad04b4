// Geometry.hpp -- problem geometry of the drop-in C++ layer (field names and
// integer types of ROCm/rocHPCG src/Geometry.hpp:29-116).  Single rank only:
// size == 1, npx = npy = npz = 1.
#ifndef HPCGX_GEOMETRY_HPP
#define HPCGX_GEOMETRY_HPP

typedef int local_int_t;         // Geometry.hpp:29
typedef long long global_int_t;  // Geometry.hpp:38

struct Geometry_STRUCT
{
    int size;
    int rank;
    int numThreads;
    local_int_t nx, ny, nz;
    int npx, npy, npz;
    int pz;
    int npartz;
    int* partz_ids;
    local_int_t* partz_nz;
    int ipx, ipy, ipz;
    global_int_t gnx, gny, gnz;
    global_int_t gix0, giy0, giz0;
};
typedef struct Geometry_STRUCT Geometry;

// GenerateGeometry.cpp:48 signature; returns nothing like the reference, and
// leaves `size`/`npx..npz` as given so OptimizeProblem can refuse multi-rank
// geometries.
void GenerateGeometry(int size, int rank, int numThreads, int pz, local_int_t zl, local_int_t zu, local_int_t nx,
                      local_int_t ny, local_int_t nz, int npx, int npy, int npz, Geometry* geom);
void DeleteGeometry(Geometry& geom);

#endif
