// xhpcg.cpp -- the HPCG benchmark run on the drop-in layer: the phases of the
// reference's main.cpp (main.cpp:183-532) -- setup, optimization, validation
// (TestCG, TestSymmetry), the optimized CG setup against the reference
// tolerance, the timed CG sets, TestNorms -- and a YAML report with the
// HPCG-3.1 field names of ReportResults.cpp:65-517.
//
// Differences from the reference's driver, all because this library has no
// host (_ref) path:
//   * the reference tolerance is the residual reduction of 50 iterations of
//     the natural-order CG_ref (main.cpp:336-345); this driver takes it from
//     --tol=, or else from 50 iterations of its own multicolour CG;
//   * "Optimization phase time vs reference SpMV+MG time" divides by this
//     library's SpMV+MG time, measured after OptimizeProblem;
//   * the per-kernel times (DDOT, WAXPBY, SpMV, MG) come from one extra CG set
//     timed phase by phase with HIP events, scaled to the timed sets' total.
//
// Usage: xhpcg [nx ny nz [rt]] [--nx=N] [--ny=N] [--nz=N] [--rt=SECONDS]
//              [--tol=T] [--quick] [--yaml=PATH]
// Without dimensions it reads hpcg.dat from the working directory, like the
// reference (ReadHpcgDat.cpp), and falls back to 16^3 and 0 s.
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <deque>
#include <string>
#include <vector>

#include "Extensions.hpp"
#include "hpcg.hpp"

namespace
{

double now()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// ---------------------------------------------------------------------------
// A small ordered YAML tree: "key: value" lines, children indented by 2.
// (A deque keeps references to earlier children valid as more are added.)
struct Yaml
{
    std::string key, value;
    std::deque<Yaml> kids;

    Yaml& add(const std::string& k, const std::string& v = "")
    {
        kids.push_back(Yaml{k, v, {}});
        return kids.back();
    }
    Yaml& add(const std::string& k, double v)
    {
        char b[64];
        snprintf(b, sizeof b, "%.6g", v);
        return add(k, std::string(b));
    }
    Yaml& add(const std::string& k, long long v) { return add(k, std::to_string(v)); }
    Yaml& add(const std::string& k, int v) { return add(k, std::to_string(v)); }
    void emit(std::string& out, int depth) const
    {
        for(const Yaml& c : kids)
        {
            out += std::string(2 * depth, ' ') + c.key + ":" + (c.value.empty() ? "" : " " + c.value) + "\n";
            c.emit(out, depth + 1);
        }
    }
};

// The run log of init.cpp:243-249: hpcgYYYYMMDDTHHMMSS.txt in the working
// directory, opened at start; main.cpp:445,506-507 write the failure count and
// one "Call [k] Scaled Residual [r]" line per timed set into it.
FILE* open_log()
{
    const time_t now_t = time(nullptr);
    const struct tm* tm = localtime(&now_t);
    char name[64];
    snprintf(name, sizeof name, "hpcg%04d%02d%02dT%02d%02d%02d.txt", 1900 + tm->tm_year, tm->tm_mon + 1,
             tm->tm_mday, tm->tm_hour, tm->tm_min, tm->tm_sec);
    FILE* f = fopen(name, "w");
    if(!f) fprintf(stderr, "xhpcg: cannot open the log file %s\n", name);
    return f;
}

// ---------------------------------------------------------------------------
struct Params
{
    int nx = 0, ny = 0, nz = 0, rt = 0, device = 0;
    bool verify = true, quick = false;
    double tol = 0.0;
    std::string yaml;
};

bool starts(const char* s, const char* p) { return strncmp(s, p, strlen(p)) == 0; }

void skip_line(FILE* f)
{
    int c;
    while((c = fgetc(f)) != EOF && c != '\n') {}
}

// ReadHpcgDat.cpp:42-73: two header lines, "nx ny nz", "seconds", then the
// optional process grid; dimensions under 16 become 16, a missing run time 1800 s.
bool read_hpcg_dat(int* dims, int* rt, int* procs)
{
    FILE* f = fopen("hpcg.dat", "r");
    if(!f) return false;
    skip_line(f);
    skip_line(f);
    for(int i = 0; i < 3; ++i)
        if(fscanf(f, "%d", dims + i) != 1 || dims[i] < 16) dims[i] = 16;
    skip_line(f);
    if(rt && (fscanf(f, "%d", rt) != 1 || *rt < 0)) *rt = 30 * 60;
    skip_line(f);
    for(int i = 0; i < 3; ++i)
        if(fscanf(f, "%d", procs + i) != 1 || procs[i] < 1) procs[i] = 0;
    fclose(f);
    return true;
}

// init.cpp:115-159: the first arguments as nx ny nz rt (values under 10 are
// ignored), "--key=" forms (--dev= picks the GPU), --tol= turns the reference CG off; without
// dimensions on the command line hpcg.dat supplies them; a dimension under 16
// takes the largest other one, or 16.
Params parse(int argc, char** argv)
{
    Params P;
    int pos[4] = {0, 0, 0, 0};
    for(int i = 0; i < 4 && i + 1 < argc; ++i)
        if(sscanf(argv[i + 1], "%d", pos + i) != 1 || pos[i] < 10) pos[i] = 0;
    P.nx = pos[0];
    P.ny = pos[1];
    P.nz = pos[2];
    int rt = pos[3];
    for(int i = 1; i < argc; ++i)
    {
        const char* a = argv[i];
        if(starts(a, "--nx="))
            P.nx = atoi(a + 5);
        else if(starts(a, "--ny="))
            P.ny = atoi(a + 5);
        else if(starts(a, "--nz="))
            P.nz = atoi(a + 5);
        else if(starts(a, "--rt="))
            rt = atoi(a + 5);
        else if(starts(a, "--dev="))
            P.device = atoi(a + 6);
        else if(starts(a, "--tol="))
        {
            P.tol = atof(a + 6);
            P.verify = false;
        }
        else if(strcmp(a, "--quick") == 0)
            P.quick = true;
        else if(starts(a, "--yaml="))
            P.yaml = a + 7;
    }
    if(!P.nx && !P.ny && !P.nz)
    {
        int d[3] = {0, 0, 0}, procs[3] = {0, 0, 0}, file_rt = 0;
        if(read_hpcg_dat(d, rt ? nullptr : &file_rt, procs))
        {
            P.nx = d[0];
            P.ny = d[1];
            P.nz = d[2];
            if(!rt) rt = file_rt;
            if(procs[0] * procs[1] * procs[2] > 1)
                fprintf(stderr, "xhpcg: hpcg.dat asks for a %d x %d x %d process grid; this run uses one GPU\n",
                        procs[0], procs[1], procs[2]);
        }
    }
    P.rt = rt;
    int d[3] = {P.nx, P.ny, P.nz};
    for(int i = 0; i < 3; ++i)
    {
        if(d[i] < 16)
            for(int j = 1; j <= 2; ++j)
                if(d[(i + j) % 3] > d[i]) d[i] = d[(i + j) % 3];
        if(d[i] < 16) d[i] = 16;
    }
    P.nx = d[0];
    P.ny = d[1];
    P.nz = d[2];
    return P;
}

// ---------------------------------------------------------------------------
// The operation and traffic models of ReportResults.cpp:65-200 for `sets`
// CG sets of `iters` iterations, 1 pre- and 1 post-smoother per level.
struct Model
{
    double ddot, waxpby, spmv, mg, total;                     // flops
    double r_ddot, w_ddot, r_wax, w_wax, r_spmv, w_spmv, r_mg, w_mg;   // bytes
};

Model model(const SparseMatrix& A, int levels, double sets, double iters)
{
    Model M{};
    const double n = sets * iters;
    const double nrow = (double)A.totalNumberOfRows, nnz = (double)A.totalNumberOfNonzeros;
    const double d = sizeof(double), ix = sizeof(local_int_t);
    M.ddot = (3.0 * n + sets) * 2.0 * nrow;
    M.waxpby = (3.0 * n + sets) * 2.0 * nrow;
    M.spmv = (n + sets) * 2.0 * nnz;
    M.r_ddot = (3.0 * n + sets) * 2.0 * nrow * d;
    M.w_ddot = (3.0 * n + sets) * d;
    M.r_wax = (3.0 * n + sets) * 2.0 * nrow * d;
    M.w_wax = (3.0 * n + sets) * nrow * d;
    M.r_spmv = (n + sets) * (nnz * (d + ix) + nrow * d);
    M.w_spmv = (n + sets) * nrow * d;
    const SparseMatrix* Af = &A;
    for(int l = 1; l < levels; ++l)
    {
        const double z = (double)Af->totalNumberOfNonzeros, r = (double)Af->totalNumberOfRows;
        M.mg += n * 4.0 * z + n * 2.0 * z + n * 4.0 * z;   // pre-smoother, residual, post-smoother
        M.r_mg += n * (2.0 * z * (d + ix) + r * d) + n * (z * (d + ix) + r * d) + n * (2.0 * z * (d + ix) + r * d);
        M.w_mg += n * r * d + n * z * d + n * z * d;
        Af = Af->Ac;
    }
    const double zc = (double)Af->totalNumberOfNonzeros, rc = (double)Af->totalNumberOfRows;
    M.mg += n * 4.0 * zc;
    M.r_mg += n * (2.0 * zc * (d + ix) + rc * d);
    M.w_mg += n * rc * d;
    M.total = M.ddot + M.waxpby + M.spmv + M.mg;
    return M;
}

// ---------------------------------------------------------------------------
// TestCG.cpp:111-177: exaggerated diagonal, 2 unpreconditioned and 2
// preconditioned CG calls to 1e-12; at most 12 / 2 iterations.
struct TestCGResult
{
    int pass = 0, fail = 0, max_no_prec = 0, max_prec = 0;
    double normr = 0.0;
};

TestCGResult test_cg(SparseMatrix& A, CGData& data, Vector& b, Vector& x)
{
    TestCGResult T;
    const local_int_t m = A.localNumberOfRows;
    Vector orig_diag, big_diag, orig_b;
    HIPInitializeVector(orig_diag, m);
    HIPInitializeVector(big_diag, m);
    HIPInitializeVector(orig_b, m);
    HIPCopyMatrixDiagonal(A, orig_diag);
    HIPCopyVector(orig_diag, big_diag);
    HIPCopyVector(b, orig_b);
    if(HPCGX_TestCGScale(A, big_diag, b)) ++T.fail;
    HIPReplaceMatrixDiagonal(A, big_diag);
    std::vector<double> times(10, 0.0);
    for(int k = 0; k < 2; ++k)
    {
        const int expected = k == 0 ? 12 : 2;
        for(int call = 0; call < 2; ++call)
        {
            HIPZeroVector(x);
            int niters = 0;
            double normr = 0.0, normr0 = 0.0;
            if(CG(A, data, b, x, 50, 1.0e-12, niters, normr, normr0, times.data(), k == 1, false)) ++T.fail;
            if(niters <= expected)
                ++T.pass;
            else
                ++T.fail;
            int& mx = k == 0 ? T.max_no_prec : T.max_prec;
            if(niters > mx) mx = niters;
            T.normr = normr;
        }
    }
    HIPReplaceMatrixDiagonal(A, orig_diag);
    HIPCopyVector(orig_b, b);
    HIPDeleteVector(orig_diag);
    HIPDeleteVector(big_diag);
    HIPDeleteVector(orig_b);
    return T;
}

// TestSymmetry.cpp:88-166: scaled departures |x'Ay - y'Ax| and
// |x'M^-1 y - y'M^-1 x| over (2 |x|^2 |A| |y|^2) eps, with |A| = 2*26;
// each above 1 counts as a failure.
struct TestSymResult
{
    double spmv = 0.0, mg = 0.0;
    double spmv_residual[2] = {-1.0, -1.0};
    int fail = 0;
};

TestSymResult test_symmetry(const SparseMatrix& A, const Vector& b, const Vector& xexact, FILE* log)
{
    TestSymResult T;
    const local_int_t n = A.localNumberOfRows;
    Vector x, y, z;
    HIPInitializeVector(x, n);
    HIPInitializeVector(y, n);
    HIPInitializeVector(z, n);
    HIPFillRandomVector(x);
    HIPFillRandomVector(y);
    double t = 0.0, xx = 0.0, yy = 0.0, xAy = 0.0, yAx = 0.0, xMy = 0.0, yMx = 0.0;
    bool opt = true;
    const double anorm = 2.0 * 26.0;
    ComputeDotProduct(n, y, y, yy, t, opt);
    ComputeSPMV(A, y, z);
    ComputeDotProduct(n, x, z, xAy, t, opt);
    ComputeDotProduct(n, x, x, xx, t, opt);
    ComputeSPMV(A, x, z);
    ComputeDotProduct(n, y, z, yAx, t, opt);
    const double scale = (xx * anorm * yy + yy * anorm * xx) * DBL_EPSILON;
    T.spmv = std::fabs(xAy - yAx) / scale;
    ComputeMG(A, y, z);
    ComputeDotProduct(n, x, z, xMy, t, opt);
    ComputeMG(A, x, z);
    ComputeDotProduct(n, y, z, yMx, t, opt);
    T.mg = std::fabs(xMy - yMx) / scale;
    T.fail = (T.spmv > 1.0) + (T.mg > 1.0);
    // TestSymmetry.cpp:150-160: A xexact against b, inf-norm, twice (logged)
    HIPCopyVector(xexact, x);
    for(int i = 0; i < 2; ++i)
    {
        ComputeSPMV(A, x, z);
        double res = -1.0;
        if(ComputeResidual(n, b, z, res)) T.fail += 1;
        T.spmv_residual[i] = res;
        if(log) fprintf(log, "SpMV call [%d] Residual [%g]\n", i, res);
    }
    HIPDeleteVector(x);
    HIPDeleteVector(y);
    HIPDeleteVector(z);
    return T;
}

// TestNorms.cpp:31-46
struct TestNormsResult
{
    double mean = 0.0, variance = 0.0;
    bool pass = false;
};

TestNormsResult test_norms(const std::vector<double>& v)
{
    TestNormsResult T;
    const double n = (double)v.size();
    double dm = 0.0;
    for(double e : v) dm += e - v[0];
    T.mean = v[0] + dm / n;
    double s = 0.0;
    for(double e : v) s += (e - T.mean) * (e - T.mean);
    T.variance = s / n;
    T.pass = T.variance < 1.0e-6;
    return T;
}

}  // namespace

int main(int argc, char** argv)
{
    setvbuf(stdout, nullptr, _IOLBF, 0);   // progress lines reach a pipe as they are printed
    const Params prm = parse(argc, argv);
    const int levels = 4;
    std::vector<double> times(10, 0.0);
    printf("xhpcg: %d x %d x %d, %d levels, run time %d s%s\n", prm.nx, prm.ny, prm.nz, levels, prm.rt,
           prm.quick ? " (quick path)" : "");
    if(prm.nx % 8 || prm.ny % 8 || prm.nz % 8)
    {
        fprintf(stderr, "xhpcg: nx, ny and nz must be multiples of 8 for %d multigrid levels\n", levels);
        return 1;
    }

    // device, stream and scratch before the setup timer, as HPCG_Init does
    if(HPCGX_Init(prm.device)) return 1;
    FILE* log = open_log();

    // ---- setup (main.cpp:183-213) --------------------------------------
    double t = now();
    Geometry* geom = new Geometry;
    GenerateGeometry(1, 0, 1, 0, 0, 0, prm.nx, prm.ny, prm.nz, 1, 1, 1, geom);
    SparseMatrix A;
    InitializeSparseMatrix(A, geom);
    Vector b, x, xexact;
    GenerateProblem(A, &b, &x, &xexact);
    SetupHalo(A);
    SparseMatrix* cur = &A;
    for(int l = 1; l < levels; ++l)
    {
        GenerateCoarseProblem(*cur);
        cur = cur->Ac;
    }
    times[9] = now() - t;
    printf("\nSetup Phase took %0.2lf sec\n", times[9]);

    // ---- optimization (main.cpp:350-361) ---------------------------------
    int global_failure = 0;
    CGData data;
    t = now();
    if(OptimizeProblem(A, data, b, x, xexact))
    {
        fprintf(stderr, "xhpcg: OptimizeProblem failed\n");
        return 1;
    }
    times[7] = now() - t;
    printf("\nOptimization Phase took %0.2lf sec\n", times[7]);

    // ---- assembled-data check (main.cpp:236-246): CheckProblem on every
    // level, b / x / xexact on the finest; the matrices exist once
    // OptimizeProblem has built them, so the check runs here, on the device
    long long check_bad = 0, check_rows = 0, check_nnz = 0;
    if(prm.verify)
    {
        printf("\nChecking assembled data ...\n");
        SparseMatrix* cm = &A;
        for(int l = 0; l < levels && cm; ++l, cm = cm->Ac)
        {
            long long c[8] = {};
            const long long bad = l == 0 ? HPCGX_CheckProblem(*cm, &b, &x, &xexact, c)
                                         : HPCGX_CheckProblem(*cm, nullptr, nullptr, nullptr, c);
            check_bad += bad < 0 ? 1 : bad;
            check_rows += c[7];
            check_nnz += c[6];
            if(bad != 0)
                printf("CheckProblem level %d: %lld mismatches (perm %lld, colouring %lld, ELL %lld, values %lld, "
                       "descriptors %lld, vectors %lld)\n",
                       l, bad, c[0], c[1], c[2], c[3], c[4], c[5]);
        }
        printf("CheckProblem: %lld rows, %lld nonzeros on %d levels, %lld mismatches\n", check_rows, check_nnz,
               levels, check_bad);
        if(check_bad)
        {
            global_failure = 1;
            if(log) fprintf(log, "CheckProblem found %lld mismatches in the assembled data.\n", check_bad);
        }
    }

    // ---- SpMV+MG timing (main.cpp:284-305, on this library's kernels) ----
    const int ncalls = prm.quick ? 1 : 10;
    {
        Vector xo, bc;
        HIPInitializeVector(xo, A.localNumberOfRows);
        HIPInitializeVector(bc, A.localNumberOfRows);
        HIPFillRandomVector(xo);
        double dot = 0.0, tdot = 0.0;
        bool opt = true;
        ComputeDotProduct(A.localNumberOfRows, xo, xo, dot, tdot, opt);   // returns after the stream drains
        t = now();
        for(int i = 0; i < ncalls; ++i)
        {
            ComputeSPMV(A, xo, bc);
            ComputeMG(A, bc, xo);
        }
        ComputeDotProduct(A.localNumberOfRows, xo, xo, dot, tdot, opt);
        times[8] = (now() - t) / ncalls;
        HIPDeleteVector(xo);
        HIPDeleteVector(bc);
    }

    // ---- reference tolerance (main.cpp:319-349) --------------------------
    const int ref_max_iters = 50;
    double ref_tol = prm.tol;
    int niters = 0;
    double normr = 0.0, normr0 = 0.0;
    if(prm.verify)
    {
        // CG_ref in natural order with the reference's unoptimized operators
        // (main.cpp:334-344), here on the device
        printf("\nReference tolerance from %d iterations of CG_ref in natural order ...\n", ref_max_iters);
        std::vector<double> rt(10, 0.0);
        if(HPCGX_ReferenceCG(A, ref_max_iters, 0.0, niters, normr, normr0, rt.data()))
        {
            global_failure = 1;
            if(log) fprintf(log, "1 error(s) in call(s) to reference CG.\n");
        }
        ref_tol = normr / normr0;
        printf("Reference CG: %d iterations in %.2f s\n", niters, rt[0]);
    }
    printf("Reference tolerance %.6e\n", ref_tol);

    // ---- validation (main.cpp:373-390) ------------------------------------
    TestCGResult tcg;
    TestSymResult tsym;
    if(prm.verify)
    {
        printf("\nValidation Testing Phase ...\n");
        tcg = test_cg(A, data, b, x);
        tsym = test_symmetry(A, b, xexact, log);
        printf("TestCG: %d passed, %d failed (max iterations %d unpreconditioned, %d preconditioned)\n", tcg.pass,
               tcg.fail, tcg.max_no_prec, tcg.max_prec);
        printf("TestSymmetry: departure SpMV %.3g, MG %.3g; SpMV residual |b - A xexact| %g, %g\n", tsym.spmv,
               tsym.mg, tsym.spmv_residual[0], tsym.spmv_residual[1]);
    }

    // ---- optimized CG setup (main.cpp:403-448) ---------------------------
    printf("\nOptimized CG Setup ...\n");
    int opt_niters = ref_max_iters;
    double worst = 0.0;
    int tol_failures = 0;
    for(int i = 0; i < ncalls; ++i)
    {
        HIPZeroVector(x);
        std::vector<double> ot(10, 0.0);
        if(CG(A, data, b, x, 10 * ref_max_iters, ref_tol, niters, normr, normr0, ot.data(), true, false))
            global_failure = 1;
        if(normr / normr0 > ref_tol) ++tol_failures;
        if(niters > opt_niters) opt_niters = niters;
        if(ot[0] > worst) worst = ot[0];
    }
    if(tol_failures)
    {
        global_failure = 1;
        printf("Failed to reduce the residual %d times.\n", tol_failures);
        if(log) fprintf(log, "Failed to reduce the residual %d times.\n", tol_failures);
    }
    printf("Optimized CG iterations per set: %d\n", opt_niters);

    // ---- timed CG sets (main.cpp:470-512) ---------------------------------
    const int nsets = (int)(prm.rt / (worst > 0.0 ? worst : 1.0)) + 1;
    printf("\nPerforming (at least) %d CG sets in %d seconds ...\n", nsets, prm.rt);
    std::vector<double> scaled;
    int sets = 0;
    while((double)prm.rt - times[0] > 0.0 || sets < nsets)
    {
        HIPZeroVector(x);
        const int ierr = CG(A, data, b, x, opt_niters, 0.0, niters, normr, normr0, times.data(), true, false);
        if(ierr)
        {
            global_failure = 1;
            if(log) fprintf(log, "Error in call to CG: %d.\n\n", ierr);
        }
        if(log) fprintf(log, "Call [%d] Scaled Residual [%g]\n", sets, normr / normr0);
        scaled.push_back(normr / normr0);
        ++sets;
        const Model M = model(A, levels, sets, opt_niters);
        const double refops = M.total * (double)ref_max_iters / opt_niters;
        printf("CG set %d / %d    %7.1lf GFlop/s    %.1lf sec left\n", sets, nsets,
               refops / (times[0] + sets * (times[7] / 10.0 + times[9] / 10.0)) / 1e9,
               prm.rt - times[0] > 0.0 ? prm.rt - times[0] : 0.0);
    }
    const TestNormsResult tn = test_norms(scaled);


    // ---- report (ReportResults.cpp:65-517) ----------------------------------
    const Model M = model(A, levels, sets, opt_niters);
    const double ratio = (double)ref_max_iters / opt_niters;
    const double refops = M.total * ratio;
    const double reads = M.r_ddot + M.r_wax + M.r_spmv + M.r_mg, writes = M.w_ddot + M.w_wax + M.w_spmv + M.w_mg;
    const double overhead_time = times[0] + sets * (times[7] / 10.0 + times[9] / 10.0);
    const double final_gflops = refops / overhead_time / 1e9;
    const double gflops24 = refops / (times[0] + sets * times[7] / 10.0) / 1e9;
    const bool valid = tcg.fail == 0 && tsym.fail == 0 && tn.pass && !global_failure;
    char devname[256] = "unknown";
    HPCGX_DeviceName(devname, sizeof devname);

    Yaml doc;
    doc.add("rochpcg_amd", "HPCG benchmark on the MI355X-native solver library");
    doc.add("Release date", "March 28, 2019");
    Yaml& ms = doc.add("Machine Summary");
    ms.add("Distributed Processes", 1);
    ms.add("Threads per processes", 1);
    ms.add("Device", std::string(devname));
    Yaml& gd = doc.add("Global Problem Dimensions");
    gd.add("Global nx", prm.nx);
    gd.add("Global ny", prm.ny);
    gd.add("Global nz", prm.nz);
    Yaml& pd = doc.add("Processor Dimensions");
    pd.add("npx", 1);
    pd.add("npy", 1);
    pd.add("npz", 1);
    Yaml& ld = doc.add("Local Domain Dimensions");
    ld.add("nx", prm.nx);
    ld.add("ny", prm.ny);
    ld.add("nz", prm.nz);
    doc.add("########## Problem Summary  ##########");
    doc.add("Setup Information").add("Setup Time", times[9]);
    Yaml& ls = doc.add("Linear System Information");
    ls.add("Number of Equations", (long long)A.totalNumberOfRows);
    ls.add("Number of Nonzero Terms", (long long)A.totalNumberOfNonzeros);
    Yaml& mgi = doc.add("Multigrid Information");
    mgi.add("Number of coarse grid levels", levels - 1);
    Yaml& cgr = mgi.add("Coarse Grids");
    const SparseMatrix* Af = &A;
    for(int l = 1; l < levels; ++l)
    {
        cgr.add("Grid Level", l);
        cgr.add("Number of Equations", (long long)Af->Ac->totalNumberOfRows);
        cgr.add("Number of Nonzero Terms", (long long)Af->Ac->totalNumberOfNonzeros);
        cgr.add("Number of Presmoother Steps", 1);
        cgr.add("Number of Postsmoother Steps", 1);
        Af = Af->Ac;
    }
    doc.add("########## Memory Use Summary  ##########");
    const double bytes = HPCGX_DeviceBytes(A);
    Yaml& mu = doc.add("Memory Use Information");
    mu.add("Total memory used for data (Gbytes)", bytes / 1e9);
    mu.add("Bytes per equation (Total memory / Number of Equations)", bytes / (double)A.totalNumberOfRows);
    doc.add("########## V&V Testing Summary  ##########");
    Yaml& sc = doc.add("Spectral Convergence Tests");
    sc.add("Result", prm.verify ? (tcg.fail == 0 ? "PASSED" : "FAILED") : "SKIPPED (--tol given)");
    Yaml& un = sc.add("Unpreconditioned");
    un.add("Maximum iteration count", tcg.max_no_prec);
    un.add("Expected iteration count", 12);
    Yaml& pr = sc.add("Preconditioned");
    pr.add("Maximum iteration count", tcg.max_prec);
    pr.add("Expected iteration count", 2);
    Yaml& ad = doc.add("Assembled Data Check");
    ad.add("Result", prm.verify ? (check_bad == 0 ? "PASSED" : "FAILED") : "SKIPPED (--tol given)");
    ad.add("Rows checked", check_rows);
    ad.add("Nonzeros checked", check_nnz);
    ad.add("Mismatches", check_bad);
    Yaml& ds = doc.add("Departure from Symmetry |x'Ay-y'Ax|/(2*||x||*||A||*||y||)/epsilon");
    ds.add("Result", prm.verify ? (tsym.fail == 0 ? "PASSED" : "FAILED") : "SKIPPED (--tol given)");
    ds.add("Departure for SpMV", tsym.spmv);
    ds.add("Departure for MG", tsym.mg);
    doc.add("########## Iterations Summary  ##########");
    Yaml& ic = doc.add("Iteration Count Information");
    ic.add("Result", global_failure ? "FAILED" : "PASSED");
    ic.add("Reference CG iterations per set", ref_max_iters);
    ic.add("Optimized CG iterations per set", opt_niters);
    ic.add("Total number of reference iterations", ref_max_iters * sets);
    ic.add("Total number of optimized iterations", opt_niters * sets);
    ic.add("Reference tolerance", ref_tol);
    ic.add("Reference tolerance source",
           prm.verify ? "50 iterations of CG_ref in natural order (device)" : "--tol");
    doc.add("########## Reproducibility Summary  ##########");
    Yaml& rp = doc.add("Reproducibility Information");
    rp.add("Result", tn.pass ? "PASSED" : "FAILED");
    rp.add("Scaled residual mean", tn.mean);
    rp.add("Scaled residual variance", tn.variance);
    doc.add("########## Performance Summary (times in sec) ##########");
    Yaml& bt = doc.add("Benchmark Time Summary");
    bt.add("Optimization phase", times[7]);
    bt.add("DDOT", times[1]);
    bt.add("WAXPBY", times[2]);
    bt.add("SpMV", times[3]);
    bt.add("MG", times[5]);
    bt.add("Total", times[0]);
    Yaml& fo = doc.add("Floating Point Operations Summary");
    fo.add("Raw DDOT", M.ddot);
    fo.add("Raw WAXPBY", M.waxpby);
    fo.add("Raw SpMV", M.spmv);
    fo.add("Raw MG", M.mg);
    fo.add("Total", M.total);
    fo.add("Total with convergence overhead", refops);
    Yaml& gb = doc.add("GB/s Summary");
    gb.add("Raw Read B/W", reads / times[0] / 1e9);
    gb.add("Raw Write B/W", writes / times[0] / 1e9);
    gb.add("Raw Total B/W", (reads + writes) / times[0] / 1e9);
    gb.add("Total with convergence and optimization phase overhead", (reads + writes) * ratio / overhead_time / 1e9);
    Yaml& gf = doc.add("GFLOP/s Summary");
    gf.add("Raw DDOT", times[1] > 0 ? M.ddot / times[1] / 1e9 : 0.0);
    gf.add("Raw WAXPBY", times[2] > 0 ? M.waxpby / times[2] / 1e9 : 0.0);
    gf.add("Raw SpMV", times[3] > 0 ? M.spmv / times[3] / 1e9 : 0.0);
    gf.add("Raw MG", times[5] > 0 ? M.mg / times[5] / 1e9 : 0.0);
    gf.add("Raw Total", M.total / times[0] / 1e9);
    gf.add("Total with convergence overhead", refops / times[0] / 1e9);
    gf.add("Total with convergence and optimization phase overhead", final_gflops);
    Yaml& uo = doc.add("User Optimization Overheads");
    uo.add("Optimization phase time (sec)", times[7]);
    uo.add("Optimization phase time vs reference SpMV+MG time", times[8] > 0 ? times[7] / times[8] : 0.0);
    Yaml& fs = doc.add("Final Summary");
    if(valid)
    {
        fs.add("HPCG result is VALID with a GFLOP/s rating of", final_gflops);
        fs.add("HPCG 2.4 rating for historical reasons is", gflops24);
        fs.add("Results are valid but execution time (sec) is", times[0]);
        if(prm.quick)
            fs.add("You have selected the QuickPath option",
                   "Results are official for legacy installed systems with confirmation from the HPCG Benchmark leaders.");
        else if(times[0] < 1800.0)
            fs.add("Official results execution time (sec) must be at least", 1800);
    }
    else
    {
        fs.add("HPCG result is", "INVALID.");
        fs.add("Please review the YAML file contents", "You may NOT submit these results for consideration.");
    }
    std::string text;
    doc.emit(text, 0);
    std::string path = prm.yaml;
    if(path.empty())
    {
        char stamp[64];
        const time_t tt = time(nullptr);
        strftime(stamp, sizeof stamp, "%Y-%m-%d_%H-%M-%S", localtime(&tt));
        path = std::string("HPCG-Benchmark_3.1_") + stamp + ".yaml";
    }
    FILE* f = fopen(path.c_str(), "w");
    if(!f || fwrite(text.data(), 1, text.size(), f) != text.size())
    {
        fprintf(stderr, "xhpcg: cannot write %s\n", path.c_str());
        if(f) fclose(f);
        return 1;
    }
    fclose(f);

    printf("\nTotal Time: %0.2lf sec\nSetup Time: %0.2lf sec\nOptimization Time: %0.2lf sec\n", times[0], times[9],
           times[7]);
    if(!valid) printf("\n*** WARNING *** INVALID RUN\n");
    printf("\nDDOT   = %7.1lf GFlop/s\nWAXPBY = %7.1lf GFlop/s\nSpMV   = %7.1lf GFlop/s\nMG     = %7.1lf GFlop/s\n",
           times[1] > 0 ? M.ddot / times[1] / 1e9 : 0.0, times[2] > 0 ? M.waxpby / times[2] / 1e9 : 0.0,
           times[3] > 0 ? M.spmv / times[3] / 1e9 : 0.0, times[5] > 0 ? M.mg / times[5] / 1e9 : 0.0);
    printf("Total  = %7.1lf GFlop/s (%7.1lf GB/s)\nFinal  = %7.1lf GFlop/s\n", M.total / times[0] / 1e9,
           (reads + writes) / times[0] / 1e9, final_gflops);
    printf("\nReport: %s\n", path.c_str());

    DeleteMatrix(A);
    HIPDeleteCGData(data);
    HIPDeleteVector(b);
    HIPDeleteVector(x);
    HIPDeleteVector(xexact);   // DeleteMatrix(A) released geom with the hierarchy
    if(log) fclose(log);
    return valid ? 0 : 2;
}
