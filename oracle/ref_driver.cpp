/*
 * ref_driver.cpp -- TEST INFRASTRUCTURE ONLY.
 *
 * A host driver around the reference's own CPU sources (the unmodified HPCG
 * "_ref" kernels of ROCm/rocHPCG, compiled where they lie by
 * oracle/build_ref.sh into oracle/_ref/libhpcg_ref.so).  It lets the tests
 * check the oracle's restatement against the reference's compiled code:
 *   - GenerateProblem_ref + SetupHalo_ref (single rank) for the natural-order
 *     problem, copied out into the oracle's row-major layout;
 *   - ComputeSPMV_ref / ComputeSYMGS_ref / ComputeMG_ref / CG_ref /
 *     ComputeDotProduct_ref / ComputeWAXPBY_ref on a hierarchy whose arrays are
 *     handed in by the caller (natural or multicolour order).
 * rocHPCG builds its coarse levels on the GPU (GenerateCoarseProblem.cpp), so
 * the MGData wiring of the hierarchy is done here on the host.
 */
#include <cstring>
#include <fstream>
#include <vector>

#include "CGData.hpp"
#include "CG_ref.hpp"
#include "ComputeDotProduct_ref.hpp"
#include "ComputeMG_ref.hpp"
#include "ComputeSPMV_ref.hpp"
#include "ComputeSYMGS_ref.hpp"
#include "ComputeWAXPBY_ref.hpp"
#include "GenerateProblem_ref.hpp"
#include "Geometry.hpp"
#include "MGData.hpp"
#include "SetupHalo_ref.hpp"
#include "SparseMatrix.hpp"
#include "Vector.hpp"
#include "hpcg.hpp"

std::ofstream HPCG_fout; /* the reference's log stream; left closed here */

static Geometry* single_rank_geometry(int nx, int ny, int nz)
{
    Geometry* g = new Geometry;
    std::memset(g, 0, sizeof(*g));
    g->size = 1;
    g->rank = 0;
    g->numThreads = 1;
    g->nx = nx;
    g->ny = ny;
    g->nz = nz;
    g->npx = g->npy = g->npz = 1;
    g->pz = 0;
    g->npartz = 1;
    g->partz_ids = new int[1];
    g->partz_ids[0] = 1;
    g->partz_nz = new local_int_t[1];
    g->partz_nz[0] = nz;
    g->gnx = nx;
    g->gny = ny;
    g->gnz = nz;
    return g;
}

extern "C" int ref_generate(int nx, int ny, int nz, int* cols, double* vals,
                            unsigned char* nnz, int* diag, double* b, double* x,
                            double* xexact)
{
    SparseMatrix A;
    InitializeSparseMatrix(A, single_rank_geometry(nx, ny, nz));
    Vector vb, vx, vxe;
    GenerateProblem_ref(A, &vb, &vx, &vxe);
    SetupHalo_ref(A);
    const local_int_t m = A.localNumberOfRows;
    for(local_int_t i = 0; i < m; ++i)
    {
        const int n = A.nonzerosInRow[i];
        nnz[i] = (unsigned char)n;
        for(int k = 0; k < 27; ++k)
        {
            cols[(size_t)i * 27 + k] = (k < n) ? A.mtxIndL[i][k] : -1;
            vals[(size_t)i * 27 + k] = (k < n) ? A.matrixValues[i][k] : 0.0;
        }
        diag[i] = (int)(A.matrixDiagonal[i] - A.matrixValues[i]);
        b[i] = vb.values[i];
        x[i] = vx.values[i];
        xexact[i] = vxe.values[i];
    }
    DeleteVector(vb);
    DeleteVector(vx);
    DeleteVector(vxe);
    /* host arrays only: release them without touching device pointers */
    delete[] A.mtxIndL[0];
    delete[] A.matrixValues[0];
    delete[] A.mtxIndG[0];
    delete[] A.nonzerosInRow;
    delete[] A.mtxIndG;
    delete[] A.mtxIndL;
    delete[] A.matrixValues;
    delete[] A.matrixDiagonal;
    DeleteGeometry(*A.geom);
    delete A.geom;
    return 0;
}

struct RefLevel
{
    SparseMatrix A;
    std::vector<local_int_t> ind;
    std::vector<double> val;
    std::vector<char> nnz;
    std::vector<local_int_t*> indp;
    std::vector<double*> valp;
    std::vector<double*> diagp;
    std::vector<local_int_t> f2c;
    MGData mg;
    Vector rc, xc, Axf;
};

struct RefHierarchy
{
    int nlevels;
    std::vector<RefLevel*> L;
};

extern "C" void* ref_wrap(int nlevels, const int* dims, const int* const* cols,
                          const double* const* vals, const unsigned char* const* nnz,
                          const int* const* diag, const int* const* f2c)
{
    RefHierarchy* h = new RefHierarchy;
    h->nlevels = nlevels;
    for(int l = 0; l < nlevels; ++l)
    {
        RefLevel* R = new RefLevel;
        const int nx = dims[3 * l], ny = dims[3 * l + 1], nz = dims[3 * l + 2];
        const local_int_t m = nx * ny * nz;
        InitializeSparseMatrix(R->A, single_rank_geometry(nx, ny, nz));
        R->ind.assign(cols[l], cols[l] + (size_t)m * 27);
        R->val.assign(vals[l], vals[l] + (size_t)m * 27);
        R->nnz.resize(m);
        R->indp.resize(m);
        R->valp.resize(m);
        R->diagp.resize(m);
        local_int_t total = 0;
        for(local_int_t i = 0; i < m; ++i)
        {
            R->nnz[i] = (char)nnz[l][i];
            R->indp[i] = R->ind.data() + (size_t)i * 27;
            R->valp[i] = R->val.data() + (size_t)i * 27;
            R->diagp[i] = R->valp[i] + diag[l][i];
            total += nnz[l][i];
        }
        R->A.totalNumberOfRows = m;
        R->A.totalNumberOfNonzeros = total;
        R->A.localNumberOfRows = m;
        R->A.localNumberOfColumns = m;
        R->A.localNumberOfNonzeros = total;
        R->A.numberOfNonzerosPerRow = 27;
        R->A.nonzerosInRow = R->nnz.data();
        R->A.mtxIndL = R->indp.data();
        R->A.matrixValues = R->valp.data();
        R->A.matrixDiagonal = R->diagp.data();
        h->L.push_back(R);
    }
    for(int l = 0; l + 1 < nlevels; ++l)
    {
        RefLevel* F = h->L[l];
        RefLevel* C = h->L[l + 1];
        const local_int_t mc = C->A.localNumberOfRows;
        F->f2c.assign(f2c[l], f2c[l] + mc);
        InitializeVector(F->rc, mc);
        InitializeVector(F->xc, mc);
        InitializeVector(F->Axf, F->A.localNumberOfColumns);
        F->mg.numberOfPresmootherSteps = 1;
        F->mg.numberOfPostsmootherSteps = 1;
        F->mg.f2cOperator = F->f2c.data();
        F->mg.rc = &F->rc;
        F->mg.xc = &F->xc;
        F->mg.Axf = &F->Axf;
        F->mg.optimizationData = 0;
        F->mg.d_f2cOperator = 0;
        F->mg.d_c2fOperator = 0;
        F->A.mgData = &F->mg;
        F->A.Ac = &C->A;
    }
    return h;
}

extern "C" void ref_destroy(void* hv)
{
    RefHierarchy* h = (RefHierarchy*)hv;
    for(RefLevel* R : h->L)
    {
        if(R->A.mgData)
        {
            DeleteVector(R->rc);
            DeleteVector(R->xc);
            DeleteVector(R->Axf);
        }
        DeleteGeometry(*R->A.geom);
        delete R->A.geom;
        delete R;
    }
    delete h;
}

static Vector view(double* p, local_int_t n)
{
    Vector v;
    v.localLength = n;
    v.values = p;
    v.optimizationData = 0;
    v.d_values = nullptr;
    return v;
}

extern "C" int ref_spmv(void* hv, int level, double* x, double* y)
{
    RefHierarchy* h = (RefHierarchy*)hv;
    SparseMatrix& A = h->L[level]->A;
    Vector vx = view(x, A.localNumberOfColumns), vy = view(y, A.localNumberOfRows);
    return ComputeSPMV_ref(A, vx, vy);
}

extern "C" int ref_symgs(void* hv, int level, double* r, double* x)
{
    RefHierarchy* h = (RefHierarchy*)hv;
    SparseMatrix& A = h->L[level]->A;
    Vector vr = view(r, A.localNumberOfRows), vx = view(x, A.localNumberOfColumns);
    return ComputeSYMGS_ref(A, vr, vx);
}

extern "C" int ref_mg(void* hv, int level, double* r, double* x)
{
    RefHierarchy* h = (RefHierarchy*)hv;
    SparseMatrix& A = h->L[level]->A;
    Vector vr = view(r, A.localNumberOfRows), vx = view(x, A.localNumberOfColumns);
    return ComputeMG_ref(A, vr, vx);
}

extern "C" double ref_dot(int n, double* x, double* y)
{
    Vector vx = view(x, n), vy = view(y, n);
    double result = 0.0, t = 0.0;
    ComputeDotProduct_ref(n, vx, (x == y) ? vx : vy, result, t);
    return result;
}

extern "C" int ref_waxpby(int n, double alpha, double* x, double beta, double* y, double* w)
{
    Vector vx = view(x, n), vy = view(y, n), vw = view(w, n);
    return ComputeWAXPBY_ref(n, alpha, vx, beta, vy, vw);
}

extern "C" int ref_cg(void* hv, double* b, double* x, int max_iter, double tolerance,
                      int precond, int* niters, double* normr, double* normr0, double* times)
{
    RefHierarchy* h = (RefHierarchy*)hv;
    SparseMatrix& A = h->L[0]->A;
    const local_int_t m = A.localNumberOfRows;
    CGData data;
    InitializeSparseCGData(A, data);
    Vector vb = view(b, m), vx = view(x, m);
    int ierr = CG_ref(A, data, vb, vx, max_iter, tolerance, *niters, *normr, *normr0, times,
                      precond != 0, false);
    DeleteCGData(data);
    return ierr;
}
