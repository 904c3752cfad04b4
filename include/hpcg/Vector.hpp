// Vector.hpp -- dense vector of the drop-in C++ layer (Vector.hpp:59-188 of
// the reference: localLength, host `values`, device `d_values`).
#ifndef HPCGX_VECTOR_HPP
#define HPCGX_VECTOR_HPP

#include <cassert>
#include <cstdlib>

#include "Geometry.hpp"

struct Vector_STRUCT
{
    local_int_t localLength;
    double* values = nullptr;          // host copy (the _ref path; optional)
    void* optimizationData = nullptr;
    double* d_values = nullptr;        // device values
};
typedef struct Vector_STRUCT Vector;

// host helpers (Vector.hpp:78-170 of the reference)
inline void InitializeVector(Vector& v, local_int_t localLength)
{
    v.localLength = localLength;
    v.values = new double[localLength];
    v.optimizationData = nullptr;
}
inline void ZeroVector(Vector& v)
{
    for(local_int_t i = 0; i < v.localLength; ++i) v.values[i] = 0.0;
}
inline void ScaleVectorValue(Vector& v, local_int_t index, double value)
{
    assert(index >= 0 && index < v.localLength);
    v.values[index] *= value;
}
inline void FillRandomVector(Vector& v)
{
    for(local_int_t i = 0; i < v.localLength; ++i) v.values[i] = rand() / (double)(RAND_MAX) + 1.0;
}
inline void CopyVector(const Vector& v, Vector& w)
{
    assert(w.localLength >= v.localLength);
    for(local_int_t i = 0; i < v.localLength; ++i) w.values[i] = v.values[i];
}
inline void DeleteVector(Vector& v)
{
    delete[] v.values;
    v.values = nullptr;
    v.localLength = 0;
}

// device helpers; asynchronous on the solver stream
void HIPInitializeVector(Vector& v, local_int_t localLength);
void HIPZeroVector(Vector& v);
void HIPFillRandomVector(Vector& v);   // rand()/RAND_MAX + 1 drawn on the host, like the reference
void HIPCopyVector(const Vector& v, Vector& w);
void HIPDeleteVector(Vector& v);

#endif
