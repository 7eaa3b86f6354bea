// Golden-vector harness for the reference CPU solver.
//
// TEST INFRASTRUCTURE ONLY.  This file is compiled (by oracle/Makefile) against
// the reference sources where they lie under /root/reference, into
// oracle/_ref/ref_harness.  It is used to write golden fixtures in this
// container (tests/golden/gen_golden.py) and, as a prebuilt binary, as the
// CPU-baseline timer in bench.py.  Nothing in the product links or calls it.
//
// The reference API has no step cap, so the run modes drive the reference's own
// public static per-step functions in loops that mirror the bodies of
//   SecondOrderSolver::solve    (Solvers/secondordersolver.cpp:288-314)
//   SecondOrderSolver::solve2D  (Solvers/secondordersolver.cpp:392-424)
//   FirstOrderSolver::solve     (Solvers/firstordersolver.cpp:435-461)
//   FirstOrderSolver::solve2D   (Solvers/firstordersolver.cpp:521-553)
// and record every time step at full precision.  With max_steps == 0 the loop
// runs to the final time exactly like the reference and, when asked, the
// reference's own solve/solve2D is called too and compared bit for bit.

#include "Solvers/secondordersolver.h"
#include "Solvers/firstordersolver.h"
#include "Solvers/mhdsecondordersolver.h"
#include "Solvers/mhdfirstordersolver.h"
#include "Solvers/hprsecondordersolver.h"
#include "Solvers/hprfirstordersolver.h"
#include "Solvers/hprforcingsolver.h"
#include "Solvers/mhdforcingsolver.h"
#include "HPR/hpracoustictensor.h"
#include "RGFM/rgfmsolver.h"
#include "Euler/eulertests.h"
#include "MHD/mhdtests.h"
#include "HPR/hprtests.h"
#include "RGFM/eulerrgfmtests.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cstdint>
#include <chrono>
#include <iostream>
#include <limits>
#include <sstream>
#include <string>
#include <vector>

namespace {

std::vector<double> readDoubles(const char *path, size_t count)
{
    std::vector<double> v(count);
    FILE *f = std::fopen(path, "rb");
    if (!f) { std::fprintf(stderr, "cannot open %s\n", path); std::exit(2); }
    size_t got = std::fread(v.data(), sizeof(double), count, f);
    std::fclose(f);
    if (got != count) { std::fprintf(stderr, "short read %s: %zu of %zu\n", path, got, count); std::exit(2); }
    return v;
}

void writeDoubles(const char *path, const std::vector<double> &v)
{
    FILE *f = std::fopen(path, "wb");
    if (!f) { std::fprintf(stderr, "cannot write %s\n", path); std::exit(2); }
    std::fwrite(v.data(), sizeof(double), v.size(), f);
    std::fclose(f);
}

EulerStateVector stateFrom(const double *p)
{
    return EulerStateVector(p[0], p[1], p[2], p[3], p[4], p[5]);
}

void putState(EulerStateVector s, double *p)
{
    p[0] = s.getDensity();
    p[1] = s.getXVelocity();
    p[2] = s.getYVelocity();
    p[3] = s.getZVelocity();
    p[4] = s.getPressure();
    p[5] = s.getReactionProgress();
}

void putVec(const std::vector<double> &v, double *p)
{
    for (size_t k = 0; k < v.size(); k++) p[k] = v[k];
}

bool sameBits(double a, double b)
{
    uint64_t x, y;
    std::memcpy(&x, &a, 8);
    std::memcpy(&y, &b, 8);
    return x == y;
}

// Silences the per-step status line the reference prints to stdout.
struct MuteCout
{
    std::ostringstream sink;
    std::streambuf *old;
    MuteCout() : old(std::cout.rdbuf(sink.rdbuf())) {}
    ~MuteCout() { std::cout.rdbuf(old); }
};

EulerMaterialParameters eulerParams(char **a)
{
    return EulerMaterialParameters(std::atof(a[0]), std::atof(a[1]), std::atof(a[2]), std::atof(a[3]), std::atof(a[4]), std::atof(a[5]));
}

// euler_run order dim ny nx dx cfl tfinal max_steps bias limiter g pinf cv ef tign krate in out_state out_dt check_api
int eulerRun(int argc, char **argv)
{
    if (argc < 22) { std::fprintf(stderr, "euler_run: too few arguments\n"); return 2; }
    int order = std::atoi(argv[2]);
    int dim = std::atoi(argv[3]);
    int ny = std::atoi(argv[4]);
    int nx = std::atoi(argv[5]);
    double dx = std::atof(argv[6]);
    double cfl = std::atof(argv[7]);
    double tfinal = std::strcmp(argv[8], "inf") == 0 ? std::numeric_limits<double>::infinity() : std::atof(argv[8]);
    int maxSteps = std::atoi(argv[9]);
    double bias = std::atof(argv[10]);
    int limiter = std::atoi(argv[11]);
    EulerMaterialParameters mp = eulerParams(argv + 12);
    const char *inPath = argv[18];
    const char *outState = argv[19];
    const char *outDt = argv[20];
    int checkApi = std::atoi(argv[21]);
    int ghost = (order == 1) ? 1 : 2;

    std::vector<double> in = readDoubles(inPath, (size_t)ny * nx * 6);
    std::vector<double> dts;
    std::vector<double> out((size_t)ny * nx * 6);
    MuteCout mute;
    auto t0 = std::chrono::steady_clock::now();

    if (dim == 1)
    {
        std::vector<EulerStateVector> cells(nx);
        for (int j = 0; j < nx; j++) cells[j] = stateFrom(&in[(size_t)j * 6]);
        std::vector<EulerStateVector> initial = cells;
        double t = 0.0;
        int it = 0;
        while (t < tfinal && (maxSteps == 0 || it < maxSteps))
        {
            std::vector<EulerStateVector> withBoundary = Solvers::insertBoundaryCells(cells, ghost);
            double dt = Solvers::computeStableTimeStep(withBoundary, dx, cfl, t, tfinal, it, mp);
            if (order == 1)
                FirstOrderSolver::computeFORCETimeStep(cells, withBoundary, dx, dt, mp);
            else
                SecondOrderSolver::computeSLICTimeStep(cells, withBoundary, dx, dt, bias, limiter, mp);
            t += dt;
            it += 1;
            dts.push_back(dt);
        }
        for (int j = 0; j < nx; j++) putState(cells[j], &out[(size_t)j * 6]);
        if (checkApi && maxSteps == 0)
        {
            std::vector<EulerStateVector> api = (order == 1) ? FirstOrderSolver::solve(initial, dx, cfl, tfinal, 0, mp)
                                                             : SecondOrderSolver::solve(initial, dx, cfl, tfinal, bias, limiter, 0, mp);
            bool match = true;
            for (int j = 0; j < nx; j++)
            {
                double p[6];
                putState(api[j], p);
                for (int f = 0; f < 6; f++) match = match && sameBits(p[f], out[(size_t)j * 6 + f]);
            }
            std::fprintf(stderr, "API_MATCH %d\n", match ? 1 : 0);
        }
    }
    else
    {
        std::vector<std::vector<EulerStateVector> > cells(ny, std::vector<EulerStateVector>(nx));
        for (int i = 0; i < ny; i++)
            for (int j = 0; j < nx; j++) cells[i][j] = stateFrom(&in[((size_t)i * nx + j) * 6]);
        std::vector<std::vector<EulerStateVector> > initial = cells;
        double t = 0.0;
        int it = 0;
        while (t < tfinal && (maxSteps == 0 || it < maxSteps))
        {
            std::vector<std::vector<EulerStateVector> > withBoundary = Solvers::insertBoundaryCells2D(cells, ghost);
            double dt = Solvers::computeStableTimeStep2D(withBoundary, dx, cfl, t, tfinal, it, mp);
            if (order == 1)
            {
                FirstOrderSolver::computeXFORCETimeStep2D(cells, withBoundary, dx, 0.5 * dt, mp);
                withBoundary = Solvers::insertBoundaryCells2D(cells, ghost);
                FirstOrderSolver::computeYFORCETimeStep2D(cells, withBoundary, dx, dt, mp);
                withBoundary = Solvers::insertBoundaryCells2D(cells, ghost);
                FirstOrderSolver::computeXFORCETimeStep2D(cells, withBoundary, dx, 0.5 * dt, mp);
            }
            else
            {
                SecondOrderSolver::computeXSLICTimeStep2D(cells, withBoundary, dx, 0.5 * dt, bias, limiter, mp);
                withBoundary = Solvers::insertBoundaryCells2D(cells, ghost);
                SecondOrderSolver::computeYSLICTimeStep2D(cells, withBoundary, dx, dt, bias, limiter, mp);
                withBoundary = Solvers::insertBoundaryCells2D(cells, ghost);
                SecondOrderSolver::computeXSLICTimeStep2D(cells, withBoundary, dx, 0.5 * dt, bias, limiter, mp);
            }
            t += dt;
            it += 1;
            dts.push_back(dt);
        }
        for (int i = 0; i < ny; i++)
            for (int j = 0; j < nx; j++) putState(cells[i][j], &out[((size_t)i * nx + j) * 6]);
        if (checkApi && maxSteps == 0)
        {
            std::vector<std::vector<EulerStateVector> > api = (order == 1) ? FirstOrderSolver::solve2D(initial, dx, cfl, tfinal, 0, mp)
                                                                           : SecondOrderSolver::solve2D(initial, dx, cfl, tfinal, bias, limiter, 0, mp);
            bool match = true;
            for (int i = 0; i < ny; i++)
                for (int j = 0; j < nx; j++)
                {
                    double p[6];
                    putState(api[i][j], p);
                    for (int f = 0; f < 6; f++) match = match && sameBits(p[f], out[((size_t)i * nx + j) * 6 + f]);
                }
            std::fprintf(stderr, "API_MATCH %d\n", match ? 1 : 0);
        }
    }
    double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::fprintf(stderr, "STEPS %zu SECONDS %.6f\n", dts.size(), secs);
    writeDoubles(outState, out);
    writeDoubles(outDt, dts);
    return 0;
}

// Per-function vectors.  Each input record is five primitive states a,b,c,d,e
// (30 doubles); each output record holds, in order:
//   slope(a,b,c) [6] | Fx(cons b) [6] | Fy(cons b) [6]
//   evolveX(a,b,c) side0, side1 [12] | evolveY(a,b,c) side0, side1 [12]
//   FORCEx(b,c) [6] | FORCEy(b,c) [6] | SLICx(a,b,c,d) [6] | SLICy(a,b,c,d) [6]
//   wave speed 1D of b [1] | wave speed 2D of b [1]
//   cons(b) [6] | prim(cons(b)) [6] | LFx(b,c) [6] | RIx(b,c) [6]
// = 92 doubles.
// euler_units n bias limiter dx ts g pinf cv ef tign krate in out
int eulerUnits(int argc, char **argv)
{
    if (argc < 15) { std::fprintf(stderr, "euler_units: too few arguments\n"); return 2; }
    int n = std::atoi(argv[2]);
    double bias = std::atof(argv[3]);
    int limiter = std::atoi(argv[4]);
    double dx = std::atof(argv[5]);
    double ts = std::atof(argv[6]);
    EulerMaterialParameters mp = eulerParams(argv + 7);
    std::vector<double> in = readDoubles(argv[13], (size_t)n * 30);
    const int rec = 92;
    std::vector<double> out((size_t)n * rec);
    for (int r = 0; r < n; r++)
    {
        const double *s = &in[(size_t)r * 30];
        EulerStateVector a = stateFrom(s), b = stateFrom(s + 6), c = stateFrom(s + 12), d = stateFrom(s + 18);
        double *o = &out[(size_t)r * rec];
        putVec(SlopeLimiters::computeSlopeVector(a, b, c, bias, limiter, mp), o); o += 6;
        putVec(EulerStateVector::computeXFluxVector(b.computeConservedVariableVector(mp), mp), o); o += 6;
        putVec(EulerStateVector::computeYFluxVector(b.computeConservedVariableVector(mp), mp), o); o += 6;
        putState(Solvers::evolveStateByHalfXTimeStep(a, b, c, dx, ts, bias, limiter, 0, mp), o); o += 6;
        putState(Solvers::evolveStateByHalfXTimeStep(a, b, c, dx, ts, bias, limiter, 1, mp), o); o += 6;
        putState(Solvers::evolveStateByHalfYTimeStep(a, b, c, dx, ts, bias, limiter, 0, mp), o); o += 6;
        putState(Solvers::evolveStateByHalfYTimeStep(a, b, c, dx, ts, bias, limiter, 1, mp), o); o += 6;
        putVec(FirstOrderSolver::computeXFORCEFlux(b, c, dx, ts, mp), o); o += 6;
        putVec(FirstOrderSolver::computeYFORCEFlux(b, c, dx, ts, mp), o); o += 6;
        putVec(SecondOrderSolver::computeXSLICFlux(a, b, c, d, dx, ts, bias, limiter, mp), o); o += 6;
        putVec(SecondOrderSolver::computeYSLICFlux(a, b, c, d, dx, ts, bias, limiter, mp), o); o += 6;
        {
            std::vector<EulerStateVector> one(1, b);
            *o++ = Solvers::computeMaximumWaveSpeed(one, mp);
            std::vector<std::vector<EulerStateVector> > one2(1, one);
            *o++ = Solvers::computeMaximumWaveSpeed2D(one2, mp);
        }
        std::vector<double> cb = b.computeConservedVariableVector(mp);
        putVec(cb, o); o += 6;
        EulerStateVector back;
        back.setConservedVariableVector(cb, mp);
        putState(back, o); o += 6;
        putVec(FirstOrderSolver::computeXLaxFriedrichsFlux(b, c, dx, ts, mp), o); o += 6;
        putVec(FirstOrderSolver::computeXRichtmyerFlux(b, c, dx, ts, mp), o); o += 6;
    }
    writeDoubles(argv[14], out);
    return 0;
}

// limiter_units n bias in out: per r -> superbee, van Leer, minbee, R coefficient, code 7 (-> minbee)
int limiterUnits(int argc, char **argv)
{
    if (argc < 6) { std::fprintf(stderr, "limiter_units: too few arguments\n"); return 2; }
    int n = std::atoi(argv[2]);
    double bias = std::atof(argv[3]);
    std::vector<double> in = readDoubles(argv[4], (size_t)n);
    std::vector<double> out((size_t)n * 5);
    for (int r = 0; r < n; r++)
    {
        out[(size_t)r * 5 + 0] = SlopeLimiters::computeSlopeLimiter(in[r], bias, 0);
        out[(size_t)r * 5 + 1] = SlopeLimiters::computeSlopeLimiter(in[r], bias, 1);
        out[(size_t)r * 5 + 2] = SlopeLimiters::computeSlopeLimiter(in[r], bias, 2);
        out[(size_t)r * 5 + 3] = SlopeLimiters::computeRCoefficient(in[r], bias);
        out[(size_t)r * 5 + 4] = SlopeLimiters::computeSlopeLimiter(in[r], bias, 7);
    }
    writeDoubles(argv[5], out);
    return 0;
}

MHDStateVector mhdStateFrom(const double *p)
{
    return MHDStateVector(p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7], p[8]);
}

void putMhdState(MHDStateVector s, double *p)
{
    std::vector<double> w = s.computePrimitiveVariableVector();
    for (int k = 0; k < 9; k++) p[k] = w[k];
}

// mhd_run order dim ny nx dx cfl tfinal max_steps bias limiter gamma pinf in out_state out_dt check_api [subcycles]
// Mirrors MHDSecondOrderSolver::solve / solve2D (mhdsecondordersolver.cpp:318-360, 501-546),
// RK4 GLM sub-cycles included (MHDForcingSolver::computeRungeKuttaTimeStep[2D]), and
// MHDFirstOrderSolver::solve / solve2D (mhdfirstordersolver.cpp:457-477, 552-578), which
// never set c_h and ignore sub-cycles.
int mhdRun(int argc, char **argv)
{
    if (argc < 18) { std::fprintf(stderr, "mhd_run: too few arguments\n"); return 2; }
    const int sub = (argc > 18) ? std::atoi(argv[18]) : 0;
    int order = std::atoi(argv[2]);
    int dim = std::atoi(argv[3]);
    int ny = std::atoi(argv[4]);
    int nx = std::atoi(argv[5]);
    double dx = std::atof(argv[6]);
    double cfl = std::atof(argv[7]);
    double tfinal = std::strcmp(argv[8], "inf") == 0 ? std::numeric_limits<double>::infinity() : std::atof(argv[8]);
    int maxSteps = std::atoi(argv[9]);
    double bias = std::atof(argv[10]);
    int limiter = std::atoi(argv[11]);
    MHDMaterialParameters mp(std::atof(argv[12]));
    mp.setStiffeningParameter(std::atof(argv[13]));
    const char *inPath = argv[14];
    const char *outState = argv[15];
    const char *outDt = argv[16];
    int checkApi = std::atoi(argv[17]);
    int ghost = (order == 1) ? 1 : 2;
    const int nf = 9;

    std::vector<double> in = readDoubles(inPath, (size_t)ny * nx * nf);
    std::vector<double> dts;
    std::vector<double> out((size_t)ny * nx * nf);
    MuteCout mute;
    auto t0 = std::chrono::steady_clock::now();
    if (dim == 1)
    {
        std::vector<MHDStateVector> cells(nx);
        for (int j = 0; j < nx; j++) cells[j] = mhdStateFrom(&in[(size_t)j * nf]);
        std::vector<MHDStateVector> initial = cells;
        MHDMaterialParameters p = mp;
        double t = 0.0;
        int it = 0;
        while (t < tfinal && (maxSteps == 0 || it < maxSteps))
        {
            std::vector<MHDStateVector> withBoundary = MHDSolvers::insertBoundaryCells(cells, ghost);
            double dt = MHDSolvers::computeStableTimeStep(withBoundary, dx, cfl, t, tfinal, it, p);
            if (order == 2)
            {
                double smax = MHDSolvers::computeMaximumWaveSpeed(withBoundary, p);
                p.setHyperbolicWaveSpeed(smax);
                p.configureParabolicDamping();
                for (int s = 0; s < sub; s++)
                {
                    withBoundary = MHDSolvers::insertBoundaryCells(cells, 1);
                    MHDForcingSolver::computeRungeKuttaTimeStep(cells, withBoundary, dx, 0.5 * (dt / sub), bias, limiter, p);
                }
            }
            withBoundary = MHDSolvers::insertBoundaryCells(cells, ghost);
            if (order == 1)
                MHDFirstOrderSolver::computeFORCETimeStep(cells, withBoundary, dx, dt, p);
            else
            {
                MHDSecondOrderSolver::computeSLICTimeStep(cells, withBoundary, dx, dt, bias, limiter, p);
                for (int s = 0; s < sub; s++)
                {
                    withBoundary = MHDSolvers::insertBoundaryCells(cells, 1);
                    MHDForcingSolver::computeRungeKuttaTimeStep(cells, withBoundary, dx, 0.5 * (dt / sub), bias, limiter, p);
                }
            }
            t += dt;
            it += 1;
            dts.push_back(dt);
        }
        for (int j = 0; j < nx; j++) putMhdState(cells[j], &out[(size_t)j * nf]);
        if (checkApi && maxSteps == 0)
        {
            std::vector<MHDStateVector> api = (order == 2)
                ? MHDSecondOrderSolver::solve(initial, dx, cfl, tfinal, bias, limiter, sub, mp)
                : MHDFirstOrderSolver::solve(initial, dx, cfl, tfinal, sub, mp);
            bool match = true;
            for (int j = 0; j < nx; j++)
            {
                double q[9];
                putMhdState(api[j], q);
                for (int f = 0; f < nf; f++) match = match && sameBits(q[f], out[(size_t)j * nf + f]);
            }
            std::fprintf(stderr, "API_MATCH %d\n", match ? 1 : 0);
        }
    }
    else
    {
        std::vector<std::vector<MHDStateVector> > cells(ny, std::vector<MHDStateVector>(nx));
        for (int i = 0; i < ny; i++)
            for (int j = 0; j < nx; j++) cells[i][j] = mhdStateFrom(&in[((size_t)i * nx + j) * nf]);
        MHDMaterialParameters p = mp;
        double t = 0.0;
        int it = 0;
        while (t < tfinal && (maxSteps == 0 || it < maxSteps))
        {
            std::vector<std::vector<MHDStateVector> > withBoundary = MHDSolvers::insertBoundaryCells2D(cells, ghost);
            double dt = MHDSolvers::computeStableTimeStep2D(withBoundary, dx, cfl, t, tfinal, it, p);
            if (order == 2)
            {
                double smax = MHDSolvers::computeMaximumWaveSpeed2D(withBoundary, p);
                p.setHyperbolicWaveSpeed(smax);
                p.configureParabolicDamping();
                for (int s = 0; s < sub; s++)
                {
                    withBoundary = MHDSolvers::insertBoundaryCells2D(cells, 1);
                    MHDForcingSolver::computeRungeKuttaTimeStep2D(cells, withBoundary, dx, 0.5 * (dt / sub), bias, limiter, p);
                }
            }
            withBoundary = MHDSolvers::insertBoundaryCells2D(cells, ghost);
            if (order == 1)
            {
                MHDFirstOrderSolver::computeXFORCETimeStep2D(cells, withBoundary, dx, 0.5 * dt, p);
                withBoundary = MHDSolvers::insertBoundaryCells2D(cells, ghost);
                MHDFirstOrderSolver::computeYFORCETimeStep2D(cells, withBoundary, dx, dt, p);
                withBoundary = MHDSolvers::insertBoundaryCells2D(cells, ghost);
                MHDFirstOrderSolver::computeXFORCETimeStep2D(cells, withBoundary, dx, 0.5 * dt, p);
            }
            else
            {
                MHDSecondOrderSolver::computeXSLICTimeStep2D(cells, withBoundary, dx, 0.5 * dt, bias, limiter, p);
                withBoundary = MHDSolvers::insertBoundaryCells2D(cells, ghost);
                MHDSecondOrderSolver::computeYSLICTimeStep2D(cells, withBoundary, dx, dt, bias, limiter, p);
                withBoundary = MHDSolvers::insertBoundaryCells2D(cells, ghost);
                MHDSecondOrderSolver::computeXSLICTimeStep2D(cells, withBoundary, dx, 0.5 * dt, bias, limiter, p);
                for (int s = 0; s < sub; s++)
                {
                    withBoundary = MHDSolvers::insertBoundaryCells2D(cells, 1);
                    MHDForcingSolver::computeRungeKuttaTimeStep2D(cells, withBoundary, dx, 0.5 * (dt / sub), bias, limiter, p);
                }
            }
            t += dt;
            it += 1;
            dts.push_back(dt);
        }
        for (int i = 0; i < ny; i++)
            for (int j = 0; j < nx; j++) putMhdState(cells[i][j], &out[((size_t)i * nx + j) * nf]);
        if (checkApi && maxSteps == 0)
        {
            std::vector<std::vector<MHDStateVector> > init2(ny, std::vector<MHDStateVector>(nx));
            for (int i = 0; i < ny; i++)
                for (int j = 0; j < nx; j++) init2[i][j] = mhdStateFrom(&in[((size_t)i * nx + j) * nf]);
            std::vector<std::vector<MHDStateVector> > api = (order == 2)
                ? MHDSecondOrderSolver::solve2D(init2, dx, cfl, tfinal, bias, limiter, sub, mp)
                : MHDFirstOrderSolver::solve2D(init2, dx, cfl, tfinal, sub, mp);
            bool match = true;
            for (int i = 0; i < ny; i++)
                for (int j = 0; j < nx; j++)
                {
                    double q[9];
                    putMhdState(api[i][j], q);
                    for (int f = 0; f < nf; f++) match = match && sameBits(q[f], out[((size_t)i * nx + j) * nf + f]);
                }
            std::fprintf(stderr, "API_MATCH %d\n", match ? 1 : 0);
        }
    }
    double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::fprintf(stderr, "STEPS %zu SECONDS %.6f\n", dts.size(), secs);
    writeDoubles(outState, out);
    writeDoubles(outDt, dts);
    return 0;
}

// mhd_units n bias limiter dx ts gamma pinf ch in out: records of five states
// (45 doubles) -> slope(9) Fx(9) Fy(9) evolveX side0/1 (18) FORCEx(9) SLICx(9)
// speed1D speed2D cons(9) prim(cons)(9) fast_x alfven = 85 doubles.
int mhdUnits(int argc, char **argv)
{
    if (argc < 12) { std::fprintf(stderr, "mhd_units: too few arguments\n"); return 2; }
    int n = std::atoi(argv[2]);
    double bias = std::atof(argv[3]);
    int limiter = std::atoi(argv[4]);
    double dx = std::atof(argv[5]);
    double ts = std::atof(argv[6]);
    MHDMaterialParameters mp(std::atof(argv[7]));
    mp.setStiffeningParameter(std::atof(argv[8]));
    mp.setHyperbolicWaveSpeed(std::atof(argv[9]));
    mp.configureParabolicDamping();
    std::vector<double> in = readDoubles(argv[10], (size_t)n * 45);
    const int rec = 85;
    std::vector<double> out((size_t)n * rec);
    for (int r = 0; r < n; r++)
    {
        const double *s = &in[(size_t)r * 45];
        MHDStateVector a = mhdStateFrom(s), b = mhdStateFrom(s + 9), c = mhdStateFrom(s + 18), d = mhdStateFrom(s + 27);
        double *o = &out[(size_t)r * rec];
        putVec(SlopeLimiters::computeSlopeVector(a, b, c, bias, limiter, mp), o); o += 9;
        putVec(MHDStateVector::computeXFluxVector(b.computeConservedVariableVector(mp), mp), o); o += 9;
        putVec(MHDStateVector::computeYFluxVector(b.computeConservedVariableVector(mp), mp), o); o += 9;
        putMhdState(MHDSolvers::evolveStateByHalfXTimeStep(a, b, c, dx, ts, bias, limiter, 0, mp), o); o += 9;
        putMhdState(MHDSolvers::evolveStateByHalfXTimeStep(a, b, c, dx, ts, bias, limiter, 1, mp), o); o += 9;
        putVec(MHDFirstOrderSolver::computeXFORCEFlux(b, c, dx, ts, mp), o); o += 9;
        putVec(MHDSecondOrderSolver::computeXSLICFlux(a, b, c, d, dx, ts, bias, limiter, mp), o); o += 9;
        {
            std::vector<MHDStateVector> one(1, b);
            *o++ = MHDSolvers::computeMaximumWaveSpeed(one, mp);
            std::vector<std::vector<MHDStateVector> > one2(1, one);
            *o++ = MHDSolvers::computeMaximumWaveSpeed2D(one2, mp);
        }
        std::vector<double> cb = b.computeConservedVariableVector(mp);
        putVec(cb, o); o += 9;
        MHDStateVector back;
        back.setConservedVariableVector(cb, mp);
        putMhdState(back, o); o += 9;
        *o++ = b.computeXFastMagnetoAcousticSpeed(mp);
        *o++ = b.computeAlfvenWaveSpeed();
    }
    writeDoubles(argv[11], out);
    return 0;
}

// HPR material parameters from argv: is_thermal rho0 p0 t_ref t0 cv gamma p_inf cs tau1 tau2 alpha
HPRMaterialParameters hprParams(char **a)
{
    HPRMaterialParameters mp("StiffenedGas", std::atoi(a[0]) != 0, false, 8.31445985, std::atof(a[1]), std::atof(a[2]), std::atof(a[3]),
                             std::atof(a[4]), std::atof(a[5]), std::atof(a[6]), std::atof(a[7]), 0.0, 0.0, 0.0, 0.0, std::atof(a[8]),
                             std::atof(a[9]), std::atof(a[10]), 1e-2, 0.75, 0.0, 1.0, std::atof(a[11]), 0.0, 0.0, 0.0);
    return mp;
}

HPRStateVector hprStateFrom(const double *p)
{
    std::vector<std::vector<double> > A(3, std::vector<double>(3));
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) A[i][j] = p[2 + 3 * i + j];
    return HPRStateVector(p[0], p[1], A, p[11], p[12], p[13], p[14], p[15], p[16]);
}

void putHprState(HPRStateVector s, double *p)
{
    p[0] = s.getDensity();
    p[1] = s.getPressure();
    std::vector<std::vector<double> > A = s.getDistortionTensor();
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) p[2 + 3 * i + j] = A[i][j];
    p[11] = s.getXVelocity();
    p[12] = s.getYVelocity();
    p[13] = s.getZVelocity();
    p[14] = s.getXThermalImpulse();
    p[15] = s.getYThermalImpulse();
    p[16] = s.getZThermalImpulse();
}

// hpr_run order dim ny nx dx cfl tfinal max_steps bias limiter subcycles <12 params> in out_state out_dt check_api
// Mirrors HPRSecondOrderSolver::solve / solve2D (hprsecondordersolver.cpp:314-352, 502-546).
int hprRun(int argc, char **argv)
{
    if (argc < 29) { std::fprintf(stderr, "hpr_run: too few arguments\n"); return 2; }
    int order = std::atoi(argv[2]);
    int dim = std::atoi(argv[3]);
    int ny = std::atoi(argv[4]);
    int nx = std::atoi(argv[5]);
    double dx = std::atof(argv[6]);
    double cfl = std::atof(argv[7]);
    double tfinal = std::strcmp(argv[8], "inf") == 0 ? std::numeric_limits<double>::infinity() : std::atof(argv[8]);
    int maxSteps = std::atoi(argv[9]);
    double bias = std::atof(argv[10]);
    int limiter = std::atoi(argv[11]);
    int sub = std::atoi(argv[12]);
    HPRMaterialParameters mp = hprParams(argv + 13);
    const char *inPath = argv[25];
    const char *outState = argv[26];
    const char *outDt = argv[27];
    int checkApi = std::atoi(argv[28]);
    const int nf = 17;
    std::vector<double> in = readDoubles(inPath, (size_t)ny * nx * nf);
    std::vector<double> dts;
    std::vector<double> out((size_t)ny * nx * nf);
    MuteCout mute;
    auto t0 = std::chrono::steady_clock::now();
    if (dim == 1)
    {
        std::vector<HPRStateVector> cells(nx);
        for (int j = 0; j < nx; j++) cells[j] = hprStateFrom(&in[(size_t)j * nf]);
        std::vector<HPRStateVector> initial = cells;
        double t = 0.0;
        int it = 0;
        while (t < tfinal && (maxSteps == 0 || it < maxSteps))
        {
            std::vector<HPRStateVector> withBoundary = HPRSolvers::insertBoundaryCells(cells, 2);
            double dt = HPRSolvers::computeStableTimeStep(withBoundary, dx, cfl, t, tfinal, it, mp);
            // HPRFirstOrderSolver::solve never sub-cycles (hprfirstordersolver.cpp:455-475)
            const int nsub = (order == 2) ? sub : 0;
            for (int s = 0; s < nsub; s++)
            {
                withBoundary = HPRSolvers::insertBoundaryCells(cells, 1);
                HPRForcingSolver::computeRungeKuttaTimeStep(cells, withBoundary, dx, 0.5 * (dt / sub), bias, limiter, mp);
            }
            withBoundary = HPRSolvers::insertBoundaryCells(cells, order == 1 ? 1 : 2);
            if (order == 1)
                HPRFirstOrderSolver::computeFORCETimeStep(cells, withBoundary, dx, dt, mp);
            else
                HPRSecondOrderSolver::computeSLICTimeStep(cells, withBoundary, dx, dt, bias, limiter, mp);
            for (int s = 0; s < nsub; s++)
            {
                withBoundary = HPRSolvers::insertBoundaryCells(cells, 1);
                HPRForcingSolver::computeRungeKuttaTimeStep(cells, withBoundary, dx, 0.5 * (dt / sub), bias, limiter, mp);
            }
            t += dt;
            it += 1;
            dts.push_back(dt);
        }
        for (int j = 0; j < nx; j++) putHprState(cells[j], &out[(size_t)j * nf]);
        if (checkApi && maxSteps == 0)
        {
            std::vector<HPRStateVector> api = (order == 2)
                ? HPRSecondOrderSolver::solve(initial, dx, cfl, tfinal, bias, limiter, sub, mp)
                : HPRFirstOrderSolver::solve(initial, dx, cfl, tfinal, sub, mp);
            bool match = true;
            for (int j = 0; j < nx; j++)
            {
                double q[17];
                putHprState(api[j], q);
                for (int f = 0; f < nf; f++) match = match && sameBits(q[f], out[(size_t)j * nf + f]);
            }
            std::fprintf(stderr, "API_MATCH %d\n", match ? 1 : 0);
        }
    }
    else
    {
        std::vector<std::vector<HPRStateVector> > cells(ny, std::vector<HPRStateVector>(nx));
        for (int i = 0; i < ny; i++)
            for (int j = 0; j < nx; j++) cells[i][j] = hprStateFrom(&in[((size_t)i * nx + j) * nf]);
        double t = 0.0;
        int it = 0;
        while (t < tfinal && (maxSteps == 0 || it < maxSteps))
        {
            std::vector<std::vector<HPRStateVector> > withBoundary = HPRSolvers::insertBoundaryCells2D(cells, 2);
            double dt = HPRSolvers::computeStableTimeStep2D(withBoundary, dx, cfl, t, tfinal, it, mp);
            for (int s = 0; s < sub; s++)
            {
                withBoundary = HPRSolvers::insertBoundaryCells2D(cells, 1);
                HPRForcingSolver::computeRungeKuttaTimeStep2D(cells, withBoundary, dx, 0.5 * (dt / sub), bias, limiter, mp);
            }
            withBoundary = HPRSolvers::insertBoundaryCells2D(cells, 2);
            HPRSecondOrderSolver::computeXSLICTimeStep2D(cells, withBoundary, dx, 0.5 * dt, bias, limiter, mp);
            withBoundary = HPRSolvers::insertBoundaryCells2D(cells, 2);
            HPRSecondOrderSolver::computeYSLICTimeStep2D(cells, withBoundary, dx, dt, bias, limiter, mp);
            withBoundary = HPRSolvers::insertBoundaryCells2D(cells, 2);
            HPRSecondOrderSolver::computeXSLICTimeStep2D(cells, withBoundary, dx, 0.5 * dt, bias, limiter, mp);
            for (int s = 0; s < sub; s++)
            {
                withBoundary = HPRSolvers::insertBoundaryCells2D(cells, 1);
                HPRForcingSolver::computeRungeKuttaTimeStep2D(cells, withBoundary, dx, 0.5 * (dt / sub), bias, limiter, mp);
            }
            t += dt;
            it += 1;
            dts.push_back(dt);
        }
        for (int i = 0; i < ny; i++)
            for (int j = 0; j < nx; j++) putHprState(cells[i][j], &out[((size_t)i * nx + j) * nf]);
    }
    double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::fprintf(stderr, "STEPS %zu SECONDS %.6f\n", dts.size(), secs);
    writeDoubles(outState, out);
    writeDoubles(outDt, dts);
    return 0;
}

// hpr_units n bias limiter dx ts <12 params> in out: records of five states (85 doubles) -> 156 doubles,
// the layout of orc_hpr_units.
int hprUnits(int argc, char **argv)
{
    if (argc < 20) { std::fprintf(stderr, "hpr_units: too few arguments\n"); return 2; }
    int n = std::atoi(argv[2]);
    double bias = std::atof(argv[3]);
    int limiter = std::atoi(argv[4]);
    double dx = std::atof(argv[5]);
    double ts = std::atof(argv[6]);
    HPRMaterialParameters mp = hprParams(argv + 7);
    std::vector<double> in = readDoubles(argv[19], (size_t)n * 85);
    const int rec = 156, nf = 17;
    std::vector<double> out((size_t)n * rec);
    for (int r = 0; r < n; r++)
    {
        const double *s = &in[(size_t)r * 85];
        HPRStateVector a = hprStateFrom(s), b = hprStateFrom(s + 17), c = hprStateFrom(s + 34), d = hprStateFrom(s + 51);
        double *o = &out[(size_t)r * rec];
        std::vector<double> cb = b.computeConservedVariableVector(mp);
        putVec(cb, o); o += nf;
        HPRStateVector back;
        back.setConservedVariableVector(cb, mp);
        putHprState(back, o);
        if (!mp.getIsThermal()) { o[14] = b.getXThermalImpulse(); o[15] = b.getYThermalImpulse(); o[16] = b.getZThermalImpulse(); }
        o += nf;
        putVec(HPRStateVector::computeXFluxVector(cb, mp), o); o += nf;
        putVec(HPRStateVector::computeYFluxVector(cb, mp), o); o += nf;
        putVec(HPRStateVector::computeSourceTermVector(cb, mp), o); o += nf;
        putHprState(HPRSolvers::evolveStateByHalfXTimeStep(a, b, c, dx, ts, bias, limiter, 0, mp), o);
        if (!mp.getIsThermal()) { o[14] = o[15] = o[16] = 0.0; }
        o += nf;
        putVec(HPRFirstOrderSolver::computeXFORCEFlux(b, c, dx, ts, mp), o); o += nf;
        putVec(HPRSecondOrderSolver::computeXSLICFlux(a, b, c, d, dx, ts, bias, limiter, mp), o); o += nf;
        {
            std::vector<double> nu = HPRForcingSolver::evolveConservedVariableVector(a.computeConservedVariableVector(mp), cb, c.computeConservedVariableVector(mp),
                                                                                    dx, ts, bias, limiter, mp);
            HPRStateVector ev;
            ev.setConservedVariableVector(nu, mp);
            putHprState(ev, o);
            if (!mp.getIsThermal()) { o[14] = o[15] = o[16] = 0.0; }
            o += nf;
        }
        *o++ = HPRAcousticTensor::computeMaximumWaveSpeed(b, mp, 0);
        *o++ = HPRAcousticTensor::computeMaximumWaveSpeed(b, mp, 1);
        {
            std::vector<std::vector<HPRStateVector> > one2(1, std::vector<HPRStateVector>(1, b));
            *o++ = HPRSolvers::computeMaximumWaveSpeed2D(one2, mp);
        }
    }
    writeDoubles(argv[20], out);
    return 0;
}

} // namespace

// rgfm_run dim ny nx dx cfl tfinal max_steps bias limiter g1 pinf1 cv1 ef1 tign1 k1 g2 pinf2 cv2 ef2 tign2 k2
//          in1 in2 phi out1 out2 out_phi out_dt check_api
// A loop over the reference's public per-step functions mirroring
// RGFMSolver::solveHLLC (rgfmsolver.cpp:745-784) and solveHLLC2D (:786-840).
int rgfmRun(int argc, char **argv)
{
    if (argc < 31) { std::fprintf(stderr, "rgfm_run: too few arguments\n"); return 2; }
    int dim = std::atoi(argv[2]);
    int ny = std::atoi(argv[3]);
    int nx = std::atoi(argv[4]);
    double dx = std::atof(argv[5]);
    double cfl = std::atof(argv[6]);
    double tfinal = std::strcmp(argv[7], "inf") == 0 ? std::numeric_limits<double>::infinity() : std::atof(argv[7]);
    int maxSteps = std::atoi(argv[8]);
    double bias = std::atof(argv[9]);
    int limiter = std::atoi(argv[10]);
    EulerMaterialParameters p1 = eulerParams(argv + 11);
    EulerMaterialParameters p2 = eulerParams(argv + 17);
    const size_t cells = (size_t)ny * nx;
    std::vector<double> in1 = readDoubles(argv[23], cells * 6);
    std::vector<double> in2 = readDoubles(argv[24], cells * 6);
    std::vector<double> phiIn = readDoubles(argv[25], cells);
    int checkApi = std::atoi(argv[30]);
    std::vector<double> out1(cells * 6), out2(cells * 6), outPhi(cells), dts;
    MuteCout mute;
    auto t0 = std::chrono::steady_clock::now();
    if (dim == 1)
    {
        std::vector<EulerStateVector> m1(nx), m2(nx);
        std::vector<double> phi(phiIn.begin(), phiIn.end());
        for (int j = 0; j < nx; j++) { m1[j] = stateFrom(&in1[(size_t)j * 6]); m2[j] = stateFrom(&in2[(size_t)j * 6]); }
        MultimaterialSystem initial(m1, m2, phi);
        double t = 0.0;
        int it = 0;
        while (t < tfinal && (maxSteps == 0 || it < maxSteps))
        {
            MultimaterialSystem g = RGFMSolver::applyRGFMBoundaryConditionsHLLC(MultimaterialSystem(m1, m2, phi), p1, p2);
            m1 = g.getMaterial1Cells();
            m2 = g.getMaterial2Cells();
            std::vector<EulerStateVector> b1 = Solvers::insertBoundaryCells(m1, 2);
            std::vector<EulerStateVector> b2 = Solvers::insertBoundaryCells(m2, 2);
            double dt = std::min(Solvers::computeStableTimeStep(b1, dx, cfl, t, tfinal, it, p1),
                                 Solvers::computeStableTimeStep(b2, dx, cfl, t, tfinal, it, p2));
            phi = RGFMSolver::updateLevelSetFunction(phi, dx, dt, m1, m2);
            SecondOrderSolver::computeSLICTimeStep(m1, b1, dx, dt, bias, limiter, p1);
            SecondOrderSolver::computeSLICTimeStep(m2, b2, dx, dt, bias, limiter, p2);
            t += dt;
            it += 1;
            dts.push_back(dt);
        }
        for (int j = 0; j < nx; j++) { putState(m1[j], &out1[(size_t)j * 6]); putState(m2[j], &out2[(size_t)j * 6]); }
        for (int j = 0; j < nx; j++) outPhi[j] = phi[j];
        if (checkApi && maxSteps == 0)
        {
            MultimaterialSystem api = RGFMSolver::solveHLLC(initial, dx, cfl, tfinal, bias, limiter, 0, 0, p1, p2);
            std::vector<EulerStateVector> a1 = api.getMaterial1Cells(), a2 = api.getMaterial2Cells();
            std::vector<double> ap = api.getLevelSetFunction();
            bool match = true;
            for (int j = 0; j < nx; j++)
            {
                double q1[6], q2[6];
                putState(a1[j], q1);
                putState(a2[j], q2);
                for (int f = 0; f < 6; f++)
                    match = match && sameBits(q1[f], out1[(size_t)j * 6 + f]) && sameBits(q2[f], out2[(size_t)j * 6 + f]);
                match = match && sameBits(ap[j], outPhi[j]);
            }
            std::fprintf(stderr, "API_MATCH %d\n", match ? 1 : 0);
        }
    }
    else
    {
        typedef std::vector<std::vector<EulerStateVector> > Grid;
        Grid m1(ny, std::vector<EulerStateVector>(nx)), m2 = m1;
        std::vector<std::vector<double> > phi(ny, std::vector<double>(nx));
        for (int i = 0; i < ny; i++)
            for (int j = 0; j < nx; j++)
            {
                m1[i][j] = stateFrom(&in1[((size_t)i * nx + j) * 6]);
                m2[i][j] = stateFrom(&in2[((size_t)i * nx + j) * 6]);
                phi[i][j] = phiIn[(size_t)i * nx + j];
            }
        MultimaterialSystem initial(m1, m2, phi);
        double t = 0.0;
        int it = 0;
        while (t < tfinal && (maxSteps == 0 || it < maxSteps))
        {
            MultimaterialSystem g = RGFMSolver::applyRGFMBoundaryConditionsHLLC2D(MultimaterialSystem(m1, m2, phi), p1, p2);
            m1 = g.getMaterial1Cells2D();
            m2 = g.getMaterial2Cells2D();
            Grid b1 = Solvers::insertBoundaryCells2D(m1, 2);
            Grid b2 = Solvers::insertBoundaryCells2D(m2, 2);
            double dt = std::min(Solvers::computeStableTimeStep2D(b1, dx, cfl, t, tfinal, it, p1),
                                 Solvers::computeStableTimeStep2D(b2, dx, cfl, t, tfinal, it, p2));
            phi = RGFMSolver::update2DLevelSetFunctionX(phi, dx, 0.5 * dt, m1, m2);
            phi = RGFMSolver::update2DLevelSetFunctionY(phi, dx, dt, m1, m2);
            phi = RGFMSolver::update2DLevelSetFunctionX(phi, dx, 0.5 * dt, m1, m2);
            SecondOrderSolver::computeXSLICTimeStep2D(m1, b1, dx, 0.5 * dt, bias, limiter, p1);
            b1 = Solvers::insertBoundaryCells2D(m1, 2);
            SecondOrderSolver::computeYSLICTimeStep2D(m1, b1, dx, dt, bias, limiter, p1);
            b1 = Solvers::insertBoundaryCells2D(m1, 2);
            SecondOrderSolver::computeXSLICTimeStep2D(m1, b1, dx, 0.5 * dt, bias, limiter, p1);
            SecondOrderSolver::computeXSLICTimeStep2D(m2, b2, dx, 0.5 * dt, bias, limiter, p2);
            b2 = Solvers::insertBoundaryCells2D(m2, 2);
            SecondOrderSolver::computeYSLICTimeStep2D(m2, b2, dx, dt, bias, limiter, p2);
            b2 = Solvers::insertBoundaryCells2D(m2, 2);
            SecondOrderSolver::computeXSLICTimeStep2D(m2, b2, dx, 0.5 * dt, bias, limiter, p2);
            t += dt;
            it += 1;
            dts.push_back(dt);
        }
        for (int i = 0; i < ny; i++)
            for (int j = 0; j < nx; j++)
            {
                putState(m1[i][j], &out1[((size_t)i * nx + j) * 6]);
                putState(m2[i][j], &out2[((size_t)i * nx + j) * 6]);
                outPhi[(size_t)i * nx + j] = phi[i][j];
            }
        if (checkApi && maxSteps == 0)
        {
            MultimaterialSystem api = RGFMSolver::solveHLLC2D(initial, dx, cfl, tfinal, bias, limiter, 0, 0, p1, p2);
            Grid a1 = api.getMaterial1Cells2D(), a2 = api.getMaterial2Cells2D();
            std::vector<std::vector<double> > ap = api.getLevelSetFunction2D();
            bool match = true;
            for (int i = 0; i < ny; i++)
                for (int j = 0; j < nx; j++)
                {
                    double q1[6], q2[6];
                    putState(a1[i][j], q1);
                    putState(a2[i][j], q2);
                    for (int f = 0; f < 6; f++)
                        match = match && sameBits(q1[f], out1[((size_t)i * nx + j) * 6 + f]) &&
                                sameBits(q2[f], out2[((size_t)i * nx + j) * 6 + f]);
                    match = match && sameBits(ap[i][j], outPhi[(size_t)i * nx + j]);
                }
            std::fprintf(stderr, "API_MATCH %d\n", match ? 1 : 0);
        }
    }
    double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    writeDoubles(argv[26], out1);
    writeDoubles(argv[27], out2);
    writeDoubles(argv[28], outPhi);
    writeDoubles(argv[29], dts);
    std::fprintf(stderr, "STEPS %zu SECONDS %.6f\n", dts.size(), secs);
    return 0;
}

// hllc_units n g1 pinf1 cv1 ef1 tign1 k1 g2 pinf2 cv2 ef2 tign2 k2 in out
// Records of two primitive states (12 doubles) -> 24 doubles:
//   solveX state (6) | left star rho | right star rho | star u | star p_x | S_L | S_R |
//   solveY state (6) | top star rho | bottom star rho | star v | star p_y | S_T | S_B
int hllcUnits(int argc, char **argv)
{
    if (argc < 17) { std::fprintf(stderr, "hllc_units: too few arguments\n"); return 2; }
    int n = std::atoi(argv[2]);
    EulerMaterialParameters p1 = eulerParams(argv + 3);
    EulerMaterialParameters p2 = eulerParams(argv + 9);
    std::vector<double> in = readDoubles(argv[15], (size_t)n * 12);
    std::vector<double> out((size_t)n * 24);
    for (int r = 0; r < n; r++)
    {
        EulerStateVector a = stateFrom(&in[(size_t)r * 12]);
        EulerStateVector b = stateFrom(&in[(size_t)r * 12 + 6]);
        double *o = &out[(size_t)r * 24];
        putState(HLLCSolver::solveX(a, b, p1, p2), o);
        o[6] = HLLCSolver::computeLeftStarRegionDensity(a, b, p1, p2);
        o[7] = HLLCSolver::computeRightStarRegionDensity(a, b, p1, p2);
        o[8] = HLLCSolver::computeStarRegionXVelocity(a, b, p1, p2);
        o[9] = HLLCSolver::computeStarRegionXPressure(a, b, p1, p2);
        o[10] = HLLCSolver::computeLeftWaveSpeed(a, b, p1, p2);
        o[11] = HLLCSolver::computeRightWaveSpeed(a, b, p1, p2);
        putState(HLLCSolver::solveY(a, b, p1, p2), o + 12);
        o[18] = HLLCSolver::computeTopStarRegionDensity(a, b, p1, p2);
        o[19] = HLLCSolver::computeBottomStarRegionDensity(a, b, p1, p2);
        o[20] = HLLCSolver::computeStarRegionYVelocity(a, b, p1, p2);
        o[21] = HLLCSolver::computeStarRegionYPressure(a, b, p1, p2);
        o[22] = HLLCSolver::computeTopWaveSpeed(a, b, p1, p2);
        o[23] = HLLCSolver::computeBottomWaveSpeed(a, b, p1, p2);
    }
    writeDoubles(argv[16], out);
    return 0;
}

// The reference's own .dat writers (EulerTests::outputSolution[2D],
// eulertests.cpp:54-108; MHDTests, mhdtests.cpp:123-160; HPRTests,
// hprtests.cpp:219-277; EulerRGFMTests, eulerrgfmtests.cpp:215-334), called on
// a state read from disk; they write their files into the working directory.
//   dat <euler|mhd|hpr|rgfm> <dim> <ny> <nx> <in> [<in2> <phi>] [hpr params x12]
int datOut(int argc, char **argv)
{
    if (argc < 7) { std::fprintf(stderr, "dat: too few arguments\n"); return 2; }
    const std::string kind = argv[2];
    const int dim = std::atoi(argv[3]), ny = std::atoi(argv[4]), nx = std::atoi(argv[5]);
    const size_t cells = (size_t)ny * nx;
    if (kind == "euler" || kind == "mhd") {
        const int nf = (kind == "euler") ? 6 : 9;
        std::vector<double> in = readDoubles(argv[6], cells * nf);
        if (kind == "euler") {
            std::vector<std::vector<EulerStateVector> > g(ny, std::vector<EulerStateVector>(nx));
            for (int i = 0; i < ny; i++)
                for (int j = 0; j < nx; j++) g[i][j] = stateFrom(&in[((size_t)i * nx + j) * nf]);
            if (dim == 1) EulerTests::outputSolution(g[0]);
            else EulerTests::outputSolution2D(g);
        } else {
            std::vector<std::vector<MHDStateVector> > g(ny, std::vector<MHDStateVector>(nx));
            for (int i = 0; i < ny; i++)
                for (int j = 0; j < nx; j++) g[i][j] = mhdStateFrom(&in[((size_t)i * nx + j) * nf]);
            if (dim == 1) MHDTests::outputSolution(g[0]);
            else MHDTests::outputSolution2D(g);
        }
        return 0;
    }
    if (kind == "hpr") {
        if (argc < 19) { std::fprintf(stderr, "dat hpr: too few arguments\n"); return 2; }
        std::vector<double> in = readDoubles(argv[6], cells * 17);
        HPRMaterialParameters mp = hprParams(argv + 7);
        std::vector<std::vector<HPRStateVector> > g(ny, std::vector<HPRStateVector>(nx));
        for (int i = 0; i < ny; i++)
            for (int j = 0; j < nx; j++) g[i][j] = hprStateFrom(&in[((size_t)i * nx + j) * 17]);
        if (dim == 1) HPRTests::outputSolution(g[0], mp);
        else HPRTests::outputSolution2D(g, mp);
        return 0;
    }
    if (kind == "rgfm") {
        if (argc < 9) { std::fprintf(stderr, "dat rgfm: too few arguments\n"); return 2; }
        std::vector<double> in1 = readDoubles(argv[6], cells * 6), in2 = readDoubles(argv[7], cells * 6);
        std::vector<double> phi = readDoubles(argv[8], cells);
        if (dim == 1) {
            std::vector<EulerStateVector> m1(nx), m2(nx);
            for (int j = 0; j < nx; j++) { m1[j] = stateFrom(&in1[(size_t)j * 6]); m2[j] = stateFrom(&in2[(size_t)j * 6]); }
            EulerRGFMTests::outputSolution(MultimaterialSystem(m1, m2, phi));
        } else {
            std::vector<std::vector<EulerStateVector> > m1(ny, std::vector<EulerStateVector>(nx)), m2 = m1;
            std::vector<std::vector<double> > ph(ny, std::vector<double>(nx));
            for (int i = 0; i < ny; i++)
                for (int j = 0; j < nx; j++) {
                    const size_t c = (size_t)i * nx + j;
                    m1[i][j] = stateFrom(&in1[c * 6]);
                    m2[i][j] = stateFrom(&in2[c * 6]);
                    ph[i][j] = phi[c];
                }
            EulerRGFMTests::outputSolution2D(MultimaterialSystem(m1, m2, ph));
        }
        return 0;
    }
    std::fprintf(stderr, "dat: unknown kind %s\n", kind.c_str());
    return 2;
}

int main(int argc, char **argv)
{
    if (argc < 2) { std::fprintf(stderr, "usage: ref_harness <mode> ...\n"); return 2; }
    std::string mode = argv[1];
    if (mode == "euler_run") return eulerRun(argc, argv);
    if (mode == "euler_units") return eulerUnits(argc, argv);
    if (mode == "limiter_units") return limiterUnits(argc, argv);
    if (mode == "mhd_run") return mhdRun(argc, argv);
    if (mode == "mhd_units") return mhdUnits(argc, argv);
    if (mode == "hpr_run") return hprRun(argc, argv);
    if (mode == "hpr_units") return hprUnits(argc, argv);
    if (mode == "rgfm_run") return rgfmRun(argc, argv);
    if (mode == "hllc_units") return hllcUnits(argc, argv);
    if (mode == "dat") return datOut(argc, argv);
    std::fprintf(stderr, "unknown mode %s\n", mode.c_str());
    return 2;
}
