// drop_in_main.cpp -- drives the reference-named C++ layer (include/hpcg) the
// way the reference's main.cpp does (main.cpp:183-532): geometry, problem,
// halo, 3 coarse levels, OptimizeProblem, then the optimized operators and CG.
// Prints one "key value" line per check; the GPU test parses them.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "Extensions.hpp"
#include "hpcg.hpp"
#include "../../include/rochpcg_amd.h"

static double norm(const Vector& v)
{
    double r = 0.0, t = 0.0;
    bool opt = true;
    ComputeDotProduct(v.localLength, v, v, r, t, opt);
    return std::sqrt(r);
}

int main(int argc, char** argv)
{
    const int nx = argc > 1 ? atoi(argv[1]) : 32;
    const int ny = argc > 2 ? atoi(argv[2]) : nx;
    const int nz = argc > 3 ? atoi(argv[3]) : nx;
    Geometry* geom = new Geometry;
    GenerateGeometry(1, 0, 1, 0, 0, 0, nx, ny, nz, 1, 1, 1, geom);
    SparseMatrix A;
    InitializeSparseMatrix(A, geom);
    Vector b, x, xexact;
    GenerateProblem(A, &b, &x, &xexact);
    SetupHalo(A);
    const int numberOfMgLevels = 4;
    SparseMatrix* cur = &A;
    for(int level = 1; level < numberOfMgLevels; ++level)
    {
        GenerateCoarseProblem(*cur);
        cur = cur->Ac;
    }
    CGData data;
    if(OptimizeProblem(A, data, b, x, xexact))
    {
        printf("optimize_failed 1\n");
        return 1;
    }
    printf("rows %d\nnonzeros %lld\nnblocks %d\nublocks %d\n", A.localNumberOfRows, A.totalNumberOfNonzeros, A.nblocks,
           A.ublocks);
    // main.cpp:236-246: CheckProblem on every level, b / x / xexact on the
    // finest (here on the optimized form, after OptimizeProblem)
    {
        long long bad = 0, rows = 0, nnz = 0;
        SparseMatrix* cm = &A;
        for(int level = 0; level < numberOfMgLevels; ++level, cm = cm->Ac)
        {
            CheckProblem(*cm, level == 0 ? &b : nullptr, level == 0 ? &x : nullptr, level == 0 ? &xexact : nullptr);
            long long c[8] = {};
            const long long e = level == 0 ? HPCGX_CheckProblem(*cm, &b, &x, &xexact, c)
                                           : HPCGX_CheckProblem(*cm, nullptr, nullptr, nullptr, c);
            bad += e < 0 ? 1 : e;
            nnz += c[6];
            rows += c[7];
        }
        printf("check_problem_mismatches %lld\ncheck_problem_rows %lld\ncheck_problem_nnz %lld\n", bad, rows, nnz);
    }
    const local_int_t m = A.localNumberOfRows;
    bool opt = true;
    double t4 = 0.0;

    // A * xexact == b (b = 27 - nnz_row, xexact = 1, both permuted)
    Vector Ax;
    HIPInitializeVector(Ax, m);
    ComputeSPMV(A, xexact, Ax);
    // TestSymmetry.cpp:150-160: ComputeResidual(b, A*xexact), the inf-norm
    double resid = -1.0;
    const int rc_resid = ComputeResidual(m, b, Ax, resid);
    printf("compute_residual %.17g\ncompute_residual_rc %d\n", resid, rc_resid);
    ComputeWAXPBY(m, 1.0, b, -1.0, Ax, Ax, opt);
    printf("spmv_residual %.17g\n", norm(Ax));

    // d_localToGlobalMap: the natural id of every local row (the identity on
    // one rank), as TestCG.cpp:129 reads it
    {
        std::vector<long long> l2g((size_t)m, -1);
        int ok = A.d_localToGlobalMap != nullptr &&
                 hpcgx_memcpy_d2h(l2g.data(), A.d_localToGlobalMap, sizeof(long long) * (size_t)m) == 0;
        for(local_int_t i = 0; ok && i < m; ++i) ok = l2g[(size_t)i] == i;
        printf("l2g_identity %d\n", ok);
    }

    // reference-mode restriction (SpMV on Axf + ComputeRestriction) equals
    // the fused one (ComputeFusedSpMVRestriction) at the injected points
    Vector r0, x0;
    HIPInitializeVector(r0, m);
    HIPInitializeVector(x0, m);
    HIPFillRandomVector(r0);
    HIPFillRandomVector(x0);
    ComputeSPMV(A, x0, *A.mgData->Axf);
    ComputeRestriction(A, r0);
    const local_int_t mc = A.Ac->localNumberOfRows;
    Vector rc1;
    HIPInitializeVector(rc1, mc);
    HIPCopyVector(*A.mgData->rc, rc1);
    ComputeFusedSpMVRestriction(A, r0, x0);
    ComputeWAXPBY(mc, 1.0, rc1, -1.0, *A.mgData->rc, rc1, opt);
    printf("restriction_diff %.17g\n", norm(rc1) / norm(*A.mgData->rc));

    // SymGS zero guess == SymGS from x = 0; prolongation adds xc at injected rows
    Vector s1, s2;
    HIPInitializeVector(s1, m);
    HIPInitializeVector(s2, m);
    HIPZeroVector(s1);
    ComputeSYMGS(A, r0, s1);
    ComputeSYMGSZeroGuess(A, r0, s2);
    ComputeWAXPBY(m, 1.0, s1, -1.0, s2, s2, opt);
    printf("symgs_zero_diff %.17g\n", norm(s2) / norm(s1));
    HIPZeroVector(s1);
    HIPFillRandomVector(*A.mgData->xc);
    ComputeProlongation(A, s1);
    printf("prolongation_norm2 %.17g\nxc_norm2 %.17g\n", norm(s1) * norm(s1),
           norm(*A.mgData->xc) * norm(*A.mgData->xc));

    // MG symmetry (TestSymmetry.cpp:111-147)
    Vector y1, z1;
    HIPInitializeVector(y1, m);
    HIPInitializeVector(z1, m);
    HIPFillRandomVector(y1);
    double xn, yn, xMy, yMx;
    ComputeDotProduct(m, x0, x0, xn, t4, opt);
    ComputeDotProduct(m, y1, y1, yn, t4, opt);
    ComputeMG(A, y1, z1);
    ComputeDotProduct(m, x0, z1, xMy, t4, opt);
    ComputeMG(A, x0, z1);
    ComputeDotProduct(m, y1, z1, yMx, t4, opt);
    printf("depsym_mg %.17g\n", std::fabs(xMy - yMx) / ((xn * 52.0 * yn + yn * 52.0 * xn) * 2.220446049250313e-16));

    // fused waxpby-dot
    double fr = 0.0;
    HIPCopyVector(y1, z1);
    ComputeFusedWAXPBYDot(m, -0.5, x0, z1, fr, t4);
    printf("fused_dot_vs_dot %.17g\n", std::fabs(fr - norm(z1) * norm(z1)) / fr);

    // CG: 50 iterations from x = 0 (the benchmark's set), then to a tolerance
    int niters = 0;
    double normr = 0.0, normr0 = 0.0;
    std::vector<double> times(10, 0.0);
    HIPZeroVector(x);
    if(CG(A, data, b, x, 50, 0.0, niters, normr, normr0, times.data(), true, argc > 4))
    {
        printf("cg_failed 1\n");
        return 1;
    }
    printf("cg50_niters %d\ncg50_scaled %.17g\n", niters, normr / normr0);
    // the per-phase timers of CG.cpp:187-198: dot, vector updates, SpMV, MG
    // (times[4], the all-reduce, stays 0 on one rank)
    printf("cg_times_phases %d\n", (int)(times[1] > 0.0 && times[2] > 0.0 && times[3] > 0.0 && times[4] == 0.0 &&
                                          times[5] > 0.0 && times[1] + times[2] + times[3] + times[5] <= times[0]));
    // CG_ref in natural order on the device (the reference tolerance)
    {
        int rn = 0;
        double rr = 0.0, rr0 = 0.0;
        std::vector<double> rt(10, 0.0);
        const int rc_ref = HPCGX_ReferenceCG(A, 50, 0.0, rn, rr, rr0, rt.data());
        printf("ref_cg_rc %d\nref_cg_niters %d\nref_cg_scaled %.17g\n", rc_ref, rn, rr / rr0);
    }
    HIPZeroVector(x);
    CG(A, data, b, x, 500, 1e-10, niters, normr, normr0, times.data(), true, false);
    printf("cg_tol_niters %d\ncg_tol_scaled %.17g\n", niters, normr / normr0);
    ComputeWAXPBY(m, 1.0, x, -1.0, xexact, x, opt);
    printf("solution_error %.17g\n", norm(x) / std::sqrt((double)m));

    // TestCG-style diagonal exaggeration through Copy/ReplaceMatrixDiagonal
    Vector d0, d1;
    HIPInitializeVector(d0, m);
    HIPInitializeVector(d1, m);
    CopyMatrixDiagonal(A, d0);
    ComputeWAXPBY(m, 1.0e6, d0, 0.0, d0, d1, opt);
    ReplaceMatrixDiagonal(A, d1);
    Vector bs;
    HIPInitializeVector(bs, m);
    ComputeWAXPBY(m, 1.0e6, b, 0.0, b, bs, opt);
    HIPZeroVector(x);
    CG(A, data, bs, x, 50, 1e-12, niters, normr, normr0, times.data(), false, false);
    printf("testcg_noprec_niters %d\n", niters);
    HIPZeroVector(x);
    CG(A, data, bs, x, 50, 1e-12, niters, normr, normr0, times.data(), true, false);
    printf("testcg_prec_niters %d\n", niters);
    ReplaceMatrixDiagonal(A, d0);

    // vectors the reference asserts against (too short, wrong length, no
    // device storage) come back as errors before any launch
    Vector shortv, nullv;
    HIPInitializeVector(shortv, m - 1);
    nullv.localLength = m;
    nullv.d_values = nullptr;
    nullv.optimizationData = nullptr;
    double dr = -1.0;
    const int bad = (ComputeSPMV(A, shortv, Ax) != 0) + (ComputeSPMV(A, x0, shortv) != 0) +
                    (ComputeSYMGS(A, r0, shortv) != 0) + (ComputeMG(A, nullv, s1) != 0) +
                    (ComputeDotProduct(m, nullv, x0, dr, t4, opt) != 0) +
                    (ComputeWAXPBY(m, 1.0, x0, 1.0, shortv, s1, opt) != 0) +
                    (CG(A, data, shortv, x, 5, 0.0, niters, normr, normr0, times.data(), true, false) != 0);
    printf("rejected_bad_vectors %d\n", bad);
    ComputeDotProduct(0, nullv, nullv, dr, t4, opt);
    printf("empty_dot %.17g\n", dr);

    for(Vector* v : {&Ax, &r0, &x0, &rc1, &s1, &s2, &y1, &z1, &d0, &d1, &bs, &shortv}) HIPDeleteVector(*v);
    DeleteMatrix(A);
    HIPDeleteCGData(data);
    HIPDeleteVector(b);
    HIPDeleteVector(x);
    HIPDeleteVector(xexact);
    printf("done 1\n");
    return 0;
}
