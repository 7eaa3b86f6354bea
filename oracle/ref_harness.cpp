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
#include "RGFM/mhdrgfmsolver.h"
#include "Euler/eulertests.h"
#include "MHD/mhdtests.h"
#include "HPR/hprtests.h"
#include "RGFM/eulerrgfmtests.h"
#include "HPR/hprmiegruneisen.h"
#include "HPR/hprsourceterms.h"
#include "Elasticity/hyperelasticmaterialparameters.h"
#include "Solvers/multiphysicssolvers.h"
#include "Euler/Multiphysics/eulermultiphysicstests.h"
#include "Euler/Multiphysics/eulerreducedtests.h"

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

// HPR material parameters as 26 doubles: the EOS code (0 StiffenedGas,
// 1 ShockMieGruneisen, 2 GodunovRomenski), is_thermal, is_plastic, then the 23
// remaining arguments of the full constructor in order (hprmaterialparameters.h:12-16).
static const char *kHprEosNames[3] = {"StiffenedGas", "ShockMieGruneisen", "GodunovRomenski"};

HPRMaterialParameters hprParamsVec(const std::vector<double> &v)
{
    return HPRMaterialParameters(kHprEosNames[(int)v[0]], v[1] != 0.0, v[2] != 0.0, v[3], v[4], v[5], v[6], v[7], v[8],
                                 v[9], v[10], v[11], v[12], v[13], v[14], v[15], v[16], v[17], v[18], v[19], v[20],
                                 v[21], v[22], v[23], v[24], v[25]);
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
int hprRunWith(HPRMaterialParameters mp, int argc, char **argv, int rest);

int hprRun(int argc, char **argv)
{
    if (argc < 29) { std::fprintf(stderr, "hpr_run: too few arguments\n"); return 2; }
    return hprRunWith(hprParams(argv + 13), argc, argv, 25);
}

// hpr_run_p: as hpr_run with the 12 parameters replaced by a file of 26 (hprParamsVec).
int hprRunP(int argc, char **argv)
{
    if (argc < 18) { std::fprintf(stderr, "hpr_run_p: too few arguments\n"); return 2; }
    return hprRunWith(hprParamsVec(readDoubles(argv[13], 26)), argc, argv, 14);
}

// argv[rest .. rest + 3]: in out_state out_dt check_api
int hprRunWith(HPRMaterialParameters mp, int argc, char **argv, int rest)
{
    (void)argc;
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
    const char *inPath = argv[rest];
    const char *outState = argv[rest + 1];
    const char *outDt = argv[rest + 2];
    int checkApi = std::atoi(argv[rest + 3]);
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
int hprUnitsWith(HPRMaterialParameters mp, int argc, char **argv, int rest);

int hprUnits(int argc, char **argv)
{
    if (argc < 20) { std::fprintf(stderr, "hpr_units: too few arguments\n"); return 2; }
    return hprUnitsWith(hprParams(argv + 7), argc, argv, 19);
}

// hpr_units_p: as hpr_units with a parameter file of 26 doubles.
int hprUnitsP(int argc, char **argv)
{
    if (argc < 10) { std::fprintf(stderr, "hpr_units_p: too few arguments\n"); return 2; }
    return hprUnitsWith(hprParamsVec(readDoubles(argv[7], 26)), argc, argv, 8);
}

// argv[rest], argv[rest + 1]: in out
int hprUnitsWith(HPRMaterialParameters mp, int argc, char **argv, int rest)
{
    (void)argc;
    int n = std::atoi(argv[2]);
    double bias = std::atof(argv[3]);
    int limiter = std::atoi(argv[4]);
    double dx = std::atof(argv[5]);
    double ts = std::atof(argv[6]);
    std::vector<double> in = readDoubles(argv[rest], (size_t)n * 85);
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
    writeDoubles(argv[rest + 1], out);
    return 0;
}

// hpr_eos_units n params in out: per record (rho, p, A row-major: 11 doubles) the
// HPRMieGruneisen functions (hprmiegruneisen.cpp:192-263) -- e(rho, p), p(rho, e),
// T(rho, p), de/drho, de/dp, dT/drho, dT/dp -- and computeTheta1Reciprocal
// (hprsourceterms.cpp:7-30): 8 doubles.  For GodunovRomenski de/drho and dT/drho
// are NaN: they call computeGruneisenCoefficientDerivative, which has no
// GodunovRomenski branch and ends without a return (hprmiegruneisen.cpp:110-125),
// undefined behaviour that crashes this build.
int hprEosUnits(int argc, char **argv)
{
    if (argc < 6) { std::fprintf(stderr, "hpr_eos_units: too few arguments\n"); return 2; }
    int n = std::atoi(argv[2]);
    HPRMaterialParameters mp = hprParamsVec(readDoubles(argv[3], 26));
    std::vector<double> in = readDoubles(argv[4], (size_t)n * 11);
    std::vector<double> out((size_t)n * 8);
    for (int r = 0; r < n; r++)
    {
        const double *s = &in[(size_t)r * 11];
        double rho = s[0], p = s[1];
        std::vector<std::vector<double> > A(3, std::vector<double>(3));
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) A[i][j] = s[2 + 3 * i + j];
        double *o = &out[(size_t)r * 8];
        o[0] = HPRMieGruneisen::computeInternalEnergy(rho, p, mp);
        o[1] = HPRMieGruneisen::computePressure(rho, o[0], mp);
        o[2] = HPRMieGruneisen::computeTemperature(rho, p, mp);
        const bool gr = mp.getEquationOfState() == "GodunovRomenski";
        o[3] = gr ? std::nan("") : HPRMieGruneisen::computeInternalEnergyDerivativeDensity(rho, p, mp);
        o[4] = HPRMieGruneisen::computeInternalEnergyDerivativePressure(rho, mp);
        o[5] = gr ? std::nan("") : HPRMieGruneisen::computeTemperatureDerivativeDensity(rho, p, mp);
        o[6] = HPRMieGruneisen::computeTemperatureDerivativePressure(rho, mp);
        o[7] = HPRSourceTerms::computeTheta1Reciprocal(rho, A, mp);
    }
    writeDoubles(argv[5], out);
    return 0;
}

// hpr_driver <barton|zhang> cellCount subcycles max_steps out_params out_in out_state out_dt:
// the initial condition and material parameters of HPRTests::solveBartonTest1 /
// solveZhangElasticPlasticTest (hprtests.cpp:59-134), built with the reference's
// own constructors (the hyperelastic entropy constructor,
// hprstatevector.cpp:34-51), then the reference's solve loop for max_steps
// steps (0: to the driver's final time).
int hprDriver(int argc, char **argv)
{
    if (argc < 10) { std::fprintf(stderr, "hpr_driver: too few arguments\n"); return 2; }
    const std::string which = argv[2];
    const int cellCount = std::atoi(argv[3]);
    std::vector<HPRStateVector> cells(cellCount);
    HPRMaterialParameters mp;
    std::vector<double> pv;
    double cellSpacing, finalTime;
    if (which == "barton")
    {
        HyperelasticMaterialParameters hp(8.9, 4.6, 2.1, 3.94 * pow(10.0, -4.0), 300.0, 1.0, 3.0, 2.0);
        pv = {2.0, 0.0, 0.0, 8.31445985, 8.9, 0.0, 300.0, 300.0, 3.94 * pow(10.0, -4.0), 1.0, 0.0, 3.909, 0.0, 0.0, 0.0,
              2.1, pow(10.0, 8.0), 0.0, 0.0, 0.75, 0.0, 1.0, 8.9 * 3.909 * sqrt(4.0 * pow(10.0, -4.0) / 300.0), 1.0, 3.0,
              2.0};
        mp = hprParamsVec(pv);
        mp.configureThermalImpulseRelaxationTime(4.0 * pow(10.0, -8.0));
        pv[17] = mp.getThermalImpulseRelaxationTime();
        cellSpacing = 1.0 / cellCount;
        finalTime = 0.06;
        std::vector<std::vector<double> > L(3, std::vector<double>(3, 0.0)), R(3, std::vector<double>(3, 0.0));
        L[0][0] = 0.98; L[1][0] = 0.02; L[1][1] = 1.0; L[1][2] = 0.1; L[2][2] = 1.0;
        R[0][0] = 1.0; R[1][1] = 1.0; R[1][2] = 0.1; R[2][2] = 1.0;
        std::vector<std::vector<double> > AL = MatrixAlgebra::computeInverseMatrix(L), AR = MatrixAlgebra::computeInverseMatrix(R);
        for (int i = 0; i < cellCount; i++)
            cells[i] = (i <= (0.5 * cellCount)) ? HPRStateVector(0.0, 0.5, 1.0, AL, 0.001, hp, mp)
                                                : HPRStateVector(0.0, 0.0, 0.0, AR, 0.0, hp, mp);
    }
    else if (which == "zhang")
    {
        HyperelasticMaterialParameters hp(2.71, 6.22, 3.16, 9.0 * pow(10.0, -4.0), 300.0, 1.0, 3.577, 2.088);
        pv = {2.0, 0.0, 1.0, 8.31445985, 2.71, 73.0, 300.0, 300.0, 9.0 * pow(10.0, -4.0), 1.0, 0.0, 5.037, 1.0, 1.338,
              0.0, 3.16, pow(10.0, -9.0), 0.0, 24.8 * pow(10.0, 9.0), 0.75, 0.2976 * pow(10.0, 9.0), 1.338, 2.0, 1.0,
              3.577, 2.088};
        mp = hprParamsVec(pv);
        cellSpacing = 10.0 / cellCount;
        finalTime = 0.6;
        std::vector<std::vector<double> > I = MatrixAlgebra::computeIdentityMatrix(3);
        for (int i = 0; i < cellCount; i++)
            cells[i] = (i <= (0.5 * cellCount)) ? HPRStateVector(0.75, 0.0, 0.0, I, 0.0, hp, mp)
                                                : HPRStateVector(-0.75, 0.0, 0.0, I, 0.0, hp, mp);
    }
    else
    {
        std::fprintf(stderr, "hpr_driver: unknown problem %s\n", which.c_str());
        return 2;
    }
    const int nf = 17;
    std::vector<double> in((size_t)cellCount * nf);
    for (int i = 0; i < cellCount; i++) putHprState(cells[i], &in[(size_t)i * nf]);
    writeDoubles(argv[6], pv);
    writeDoubles(argv[7], in);
    std::fprintf(stderr, "DX %.17g TFINAL %.17g\n", cellSpacing, finalTime);
    // max_steps < 0: the initial condition only.  The GodunovRomenski wave speed
    // reaches the undefined computeGruneisenCoefficientDerivative branch (see
    // hpr_eos_units), so the reference cannot run these two drivers.
    if (std::atoi(argv[5]) < 0) return 0;
    // the driver's solve (CFL 0.8, superbee, no bias) as hpr_run's 1D loop
    char dxs[64], ts[64], ms[32];
    std::snprintf(dxs, sizeof(dxs), "%.17g", cellSpacing);
    std::snprintf(ts, sizeof(ts), "%.17g", finalTime);
    std::snprintf(ms, sizeof(ms), "%d", std::atoi(argv[5]));
    char a0[] = "hpr_run", a1[] = "2", a2[] = "1", a3[] = "1", cfl[] = "0.8", bias[] = "0", lim[] = "0", chk[] = "0";
    char ncell[32];
    std::snprintf(ncell, sizeof(ncell), "%d", cellCount);
    char *args[] = {argv[0], a0, a1, a2, a3, ncell, dxs, cfl, ts, ms, bias, lim, argv[4], argv[7], argv[8], argv[9], chk};
    return hprRunWith(mp, 17, args, 13);
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


// ---- two-material diffuse-interface Euler models ------------------------------
// State vectors are built from and read back to primitive records through the
// reference's own setPrimitiveVariableVector / computePrimitiveVariableVector.
template <class S>
S mpFrom(const double *p, int np)
{
    S s;
    s.setPrimitiveVariableVector(std::vector<double>(p, p + np));
    return s;
}

template <class S>
void mpPut(S s, double *p)
{
    putVec(s.computePrimitiveVariableVector(), p);
}

// mp_run kind order dim ny nx dx cfl tfinal max_steps bias limiter reinit
//        m1(g pinf cv ef tign krate) m2(6) in out_state out_dt check_api
// The loops mirror SecondOrderSolver::solve[2D] for these states
// (secondordersolver.cpp:316-390, 426-517: reinitialisation after every
// reinit-th step) and FirstOrderSolver::solve[2D] (firstordersolver.cpp:463-640:
// none), with a step cap; check_api compares with the reference's own solve.
template <class S>
int mpRunT(int argc, char **argv, int np)
{
    int order = std::atoi(argv[3]);
    int dim = std::atoi(argv[4]);
    int ny = std::atoi(argv[5]);
    int nx = std::atoi(argv[6]);
    double dx = std::atof(argv[7]);
    double cfl = std::atof(argv[8]);
    double tfinal = std::strcmp(argv[9], "inf") == 0 ? std::numeric_limits<double>::infinity() : std::atof(argv[9]);
    int maxSteps = std::atoi(argv[10]);
    double bias = std::atof(argv[11]);
    int limiter = std::atoi(argv[12]);
    int reinit = std::atoi(argv[13]);
    EulerMaterialParameters m1 = eulerParams(argv + 14);
    EulerMaterialParameters m2 = eulerParams(argv + 20);
    const char *inPath = argv[26];
    const char *outState = argv[27];
    const char *outDt = argv[28];
    int checkApi = std::atoi(argv[29]);
    int ghost = (order == 1) ? 1 : 2;
    std::vector<double> in = readDoubles(inPath, (size_t)ny * nx * np);
    std::vector<double> out((size_t)ny * nx * np);
    std::vector<double> dts;
    MuteCout mute;
    if (dim == 1)
    {
        std::vector<S> cells(nx);
        for (int j = 0; j < nx; j++) cells[j] = mpFrom<S>(&in[(size_t)j * np], np);
        std::vector<S> initial = cells;
        double t = 0.0;
        int it = 0;
        while (t < tfinal && (maxSteps == 0 || it < maxSteps))
        {
            std::vector<S> wb = Solvers::insertBoundaryCells(cells, ghost);
            double dt = Solvers::computeStableTimeStep(wb, dx, cfl, t, tfinal, it, m1, m2);
            if (order == 1)
                FirstOrderSolver::computeFORCETimeStep(cells, wb, dx, dt, m1, m2);
            else
                SecondOrderSolver::computeSLICTimeStep(cells, wb, dx, dt, bias, limiter, m1, m2);
            t += dt;
            it += 1;
            dts.push_back(dt);
            if (order == 2 && reinit != 0 && (it % reinit) == 0) MultiphysicsSolvers::reinitialiseVolumeFraction(cells, m1, m2);
        }
        for (int j = 0; j < nx; j++) mpPut(cells[j], &out[(size_t)j * np]);
        if (checkApi && maxSteps == 0)
        {
            std::vector<S> api = (order == 1) ? FirstOrderSolver::solve(initial, dx, cfl, tfinal, 0, m1, m2)
                                              : SecondOrderSolver::solve(initial, dx, cfl, tfinal, bias, limiter, 0, reinit, m1, m2);
            bool match = true;
            std::vector<double> p(np);
            for (int j = 0; j < nx; j++)
            {
                mpPut(api[j], p.data());
                for (int f = 0; f < np; f++) match = match && sameBits(p[f], out[(size_t)j * np + f]);
            }
            std::fprintf(stderr, "API_MATCH %d\n", match ? 1 : 0);
        }
    }
    else
    {
        std::vector<std::vector<S> > cells(ny, std::vector<S>(nx));
        for (int i = 0; i < ny; i++)
            for (int j = 0; j < nx; j++) cells[i][j] = mpFrom<S>(&in[((size_t)i * nx + j) * np], np);
        std::vector<std::vector<S> > initial = cells;
        double t = 0.0;
        int it = 0;
        while (t < tfinal && (maxSteps == 0 || it < maxSteps))
        {
            std::vector<std::vector<S> > wb = Solvers::insertBoundaryCells2D(cells, ghost);
            double dt = Solvers::computeStableTimeStep2D(wb, dx, cfl, t, tfinal, it, m1, m2);
            if (order == 1)
            {
                FirstOrderSolver::computeXFORCETimeStep2D(cells, wb, dx, 0.5 * dt, m1, m2);
                wb = Solvers::insertBoundaryCells2D(cells, ghost);
                FirstOrderSolver::computeYFORCETimeStep2D(cells, wb, dx, dt, m1, m2);
                wb = Solvers::insertBoundaryCells2D(cells, ghost);
                FirstOrderSolver::computeXFORCETimeStep2D(cells, wb, dx, 0.5 * dt, m1, m2);
            }
            else
            {
                SecondOrderSolver::computeXSLICTimeStep2D(cells, wb, dx, 0.5 * dt, bias, limiter, m1, m2);
                wb = Solvers::insertBoundaryCells2D(cells, ghost);
                SecondOrderSolver::computeYSLICTimeStep2D(cells, wb, dx, dt, bias, limiter, m1, m2);
                wb = Solvers::insertBoundaryCells2D(cells, ghost);
                SecondOrderSolver::computeXSLICTimeStep2D(cells, wb, dx, 0.5 * dt, bias, limiter, m1, m2);
            }
            t += dt;
            it += 1;
            dts.push_back(dt);
            if (order == 2 && reinit != 0 && (it % reinit) == 0) MultiphysicsSolvers::reinitialiseVolumeFraction(cells, m1, m2);
        }
        for (int i = 0; i < ny; i++)
            for (int j = 0; j < nx; j++) mpPut(cells[i][j], &out[((size_t)i * nx + j) * np]);
        if (checkApi && maxSteps == 0)
        {
            std::vector<std::vector<S> > api = (order == 1) ? FirstOrderSolver::solve2D(initial, dx, cfl, tfinal, 0, m1, m2)
                                                            : SecondOrderSolver::solve2D(initial, dx, cfl, tfinal, bias, limiter, 0, reinit, m1, m2);
            bool match = true;
            std::vector<double> p(np);
            for (int i = 0; i < ny; i++)
                for (int j = 0; j < nx; j++)
                {
                    mpPut(api[i][j], p.data());
                    for (int f = 0; f < np; f++) match = match && sameBits(p[f], out[((size_t)i * nx + j) * np + f]);
                }
            std::fprintf(stderr, "API_MATCH %d\n", match ? 1 : 0);
        }
    }
    std::fprintf(stderr, "STEPS %zu\n", dts.size());
    writeDoubles(outState, out);
    writeDoubles(outDt, dts);
    return 0;
}

int mpRun(int argc, char **argv)
{
    if (argc < 30) { std::fprintf(stderr, "mp_run: too few arguments\n"); return 2; }
    return std::atoi(argv[2]) == 0 ? mpRunT<EulerMultiphysicsStateVector>(argc, argv, 8)
                                   : mpRunT<EulerReducedStateVector>(argc, argv, 7);
}

// mp_units kind n bias limiter dx ts m1(6) m2(6) in out: the record layout of
// orc_mp_units (four primitive states in; cons, prim(cons), Fx, Fy, wave speeds
// 1D/2D, slope, evolveX sides 0/1, FORCEx, SLICx, SLICy out).
template <class S>
int mpUnitsT(int argc, char **argv, int np, int nf)
{
    int n = std::atoi(argv[3]);
    double bias = std::atof(argv[4]);
    int limiter = std::atoi(argv[5]);
    double dx = std::atof(argv[6]);
    double ts = std::atof(argv[7]);
    EulerMaterialParameters m1 = eulerParams(argv + 8);
    EulerMaterialParameters m2 = eulerParams(argv + 14);
    std::vector<double> in = readDoubles(argv[20], (size_t)n * 4 * np);
    const int rec = 7 * nf + 3 * np + 2;
    std::vector<double> out((size_t)n * rec);
    for (int r = 0; r < n; r++)
    {
        const double *q = &in[(size_t)r * 4 * np];
        S a = mpFrom<S>(q, np), b = mpFrom<S>(q + np, np), c = mpFrom<S>(q + 2 * np, np), d = mpFrom<S>(q + 3 * np, np);
        double *o = &out[(size_t)r * rec];
        std::vector<double> U = S(b).computeConservedVariableVector(m1, m2);
        putVec(U, o);
        o += nf;
        S back;
        back.setConservedVariableVector(U, m1, m2);
        mpPut(back, o);
        o += np;
        putVec(S::computeXFluxVector(U, m1, m2), o);
        o += nf;
        putVec(S::computeYFluxVector(U, m1, m2), o);
        o += nf;
        std::vector<S> one(1, b);
        std::vector<std::vector<S> > one2(1, std::vector<S>(1, b));
        *o++ = Solvers::computeMaximumWaveSpeed(one, m1, m2);
        *o++ = Solvers::computeMaximumWaveSpeed2D(one2, m1, m2);
        putVec(SlopeLimiters::computeSlopeVector(a, b, c, bias, limiter, m1, m2), o);
        o += nf;
        mpPut(Solvers::evolveStateByHalfXTimeStep(a, b, c, dx, ts, bias, limiter, 0, m1, m2), o);
        o += np;
        mpPut(Solvers::evolveStateByHalfXTimeStep(a, b, c, dx, ts, bias, limiter, 1, m1, m2), o);
        o += np;
        putVec(FirstOrderSolver::computeXFORCEFlux(b, c, dx, ts, m1, m2), o);
        o += nf;
        putVec(SecondOrderSolver::computeXSLICFlux(a, b, c, d, dx, ts, bias, limiter, m1, m2), o);
        o += nf;
        putVec(SecondOrderSolver::computeYSLICFlux(a, b, c, d, dx, ts, bias, limiter, m1, m2), o);
    }
    writeDoubles(argv[21], out);
    return 0;
}

int mpUnits(int argc, char **argv)
{
    if (argc < 22) { std::fprintf(stderr, "mp_units: too few arguments\n"); return 2; }
    return std::atoi(argv[2]) == 0 ? mpUnitsT<EulerMultiphysicsStateVector>(argc, argv, 8, 9)
                                   : mpUnitsT<EulerReducedStateVector>(argc, argv, 7, 8);
}

// mp_reinit kind dim ny nx m1(6) m2(6) in out: one
// MultiphysicsSolvers::reinitialiseVolumeFraction on the grid.
template <class S>
int mpReinitT(int argc, char **argv, int np)
{
    int dim = std::atoi(argv[3]);
    int ny = std::atoi(argv[4]);
    int nx = std::atoi(argv[5]);
    EulerMaterialParameters m1 = eulerParams(argv + 6);
    EulerMaterialParameters m2 = eulerParams(argv + 12);
    std::vector<double> in = readDoubles(argv[18], (size_t)ny * nx * np);
    std::vector<double> out((size_t)ny * nx * np);
    if (dim == 1)
    {
        std::vector<S> cells(nx);
        for (int j = 0; j < nx; j++) cells[j] = mpFrom<S>(&in[(size_t)j * np], np);
        MultiphysicsSolvers::reinitialiseVolumeFraction(cells, m1, m2);
        for (int j = 0; j < nx; j++) mpPut(cells[j], &out[(size_t)j * np]);
    }
    else
    {
        std::vector<std::vector<S> > cells(ny, std::vector<S>(nx));
        for (int i = 0; i < ny; i++)
            for (int j = 0; j < nx; j++) cells[i][j] = mpFrom<S>(&in[((size_t)i * nx + j) * np], np);
        MultiphysicsSolvers::reinitialiseVolumeFraction(cells, m1, m2);
        for (int i = 0; i < ny; i++)
            for (int j = 0; j < nx; j++) mpPut(cells[i][j], &out[((size_t)i * nx + j) * np]);
    }
    writeDoubles(argv[19], out);
    return 0;
}

int mpReinit(int argc, char **argv)
{
    if (argc < 20) { std::fprintf(stderr, "mp_reinit: too few arguments\n"); return 2; }
    return std::atoi(argv[2]) == 0 ? mpReinitT<EulerMultiphysicsStateVector>(argc, argv, 8)
                                   : mpReinitT<EulerReducedStateVector>(argc, argv, 7);
}


// ---- two-material MHD RGFM -----------------------------------------------------
MHDMaterialParameters mhdParams4(char **a)
{
    MHDMaterialParameters m(std::atof(a[0]));
    m.setStiffeningParameter(std::atof(a[1]));
    m.setHyperbolicWaveSpeed(std::atof(a[2]));
    m.setParabolicDamping(std::atof(a[3]));
    return m;
}

// mhd_rgfm_run dim ny nx dx cfl tfinal max_steps bias limiter subcycles
//              p1(g pinf ch cp) p2(4) m1_in m2_in phi_in m1_out m2_out phi_out dt_out check_api
// The loops mirror MHDRGFMSolver::solve (mhdrgfmsolver.cpp:522-563) and
// solve2D (:565-693) through the reference's own public static steps, with a
// step cap; check_api compares with solve / solve2D themselves.
int mhdRgfmRun(int argc, char **argv)
{
    if (argc < 28) { std::fprintf(stderr, "mhd_rgfm_run: too few arguments\n"); return 2; }
    int dim = std::atoi(argv[2]);
    int ny = std::atoi(argv[3]);
    int nx = std::atoi(argv[4]);
    double dx = std::atof(argv[5]);
    double cfl = std::atof(argv[6]);
    double tfinal = std::strcmp(argv[7], "inf") == 0 ? std::numeric_limits<double>::infinity() : std::atof(argv[7]);
    int maxSteps = std::atoi(argv[8]);
    double bias = std::atof(argv[9]);
    int limiter = std::atoi(argv[10]);
    int sub = std::atoi(argv[11]);
    MHDMaterialParameters p1 = mhdParams4(argv + 12);
    MHDMaterialParameters p2 = mhdParams4(argv + 16);
    const size_t cells = (size_t)ny * nx;
    std::vector<double> in1 = readDoubles(argv[20], cells * 9);
    std::vector<double> in2 = readDoubles(argv[21], cells * 9);
    std::vector<double> phiIn = readDoubles(argv[22], cells);
    int checkApi = std::atoi(argv[27]);
    std::vector<double> out1(cells * 9), out2(cells * 9), outPhi(cells), dts;
    MuteCout mute;
    bool match = true;
    if (dim == 1)
    {
        std::vector<MHDStateVector> m1(nx), m2(nx);
        std::vector<double> phi(phiIn.begin(), phiIn.end());
        for (int j = 0; j < nx; j++) { m1[j] = mhdStateFrom(&in1[(size_t)j * 9]); m2[j] = mhdStateFrom(&in2[(size_t)j * 9]); }
        MHDMultimaterialSystem initial(m1, m2, phi);
        double t = 0.0;
        int it = 0;
        while (t < tfinal && (maxSteps == 0 || it < maxSteps))
        {
            MHDMultimaterialSystem g = MHDRGFMSolver::applyRGFMBoundaryConditions(MHDMultimaterialSystem(m1, m2, phi), p1, p2);
            m1 = g.getMaterial1Cells();
            m2 = g.getMaterial2Cells();
            std::vector<MHDStateVector> b1 = MHDSolvers::insertBoundaryCells(m1, 2);
            std::vector<MHDStateVector> b2 = MHDSolvers::insertBoundaryCells(m2, 2);
            double dt = std::min(MHDSolvers::computeStableTimeStep(b1, dx, cfl, t, tfinal, it, p1),
                                 MHDSolvers::computeStableTimeStep(b2, dx, cfl, t, tfinal, it, p2));
            phi = MHDRGFMSolver::updateLevelSetFunction(phi, dx, dt, m1, m2);
            MHDSecondOrderSolver::computeSLICTimeStep(m1, b1, dx, dt, bias, limiter, p1);
            MHDSecondOrderSolver::computeSLICTimeStep(m2, b2, dx, dt, bias, limiter, p2);
            t += dt;
            it += 1;
            dts.push_back(dt);
        }
        for (int j = 0; j < nx; j++) { putMhdState(m1[j], &out1[(size_t)j * 9]); putMhdState(m2[j], &out2[(size_t)j * 9]); }
        for (int j = 0; j < nx; j++) outPhi[j] = phi[j];
        if (checkApi && maxSteps == 0)
        {
            MHDMultimaterialSystem api = MHDRGFMSolver::solve(initial, dx, cfl, tfinal, bias, limiter, sub, 0, p1, p2);
            std::vector<MHDStateVector> a1 = api.getMaterial1Cells(), a2 = api.getMaterial2Cells();
            std::vector<double> ap = api.getLevelSetFunction();
            for (int j = 0; j < nx; j++)
            {
                double q1[9], q2[9];
                putMhdState(a1[j], q1);
                putMhdState(a2[j], q2);
                for (int f = 0; f < 9; f++)
                    match = match && sameBits(q1[f], out1[(size_t)j * 9 + f]) && sameBits(q2[f], out2[(size_t)j * 9 + f]);
                match = match && sameBits(ap[j], outPhi[j]);
            }
        }
    }
    else
    {
        typedef std::vector<std::vector<MHDStateVector> > Grid;
        Grid m1(ny, std::vector<MHDStateVector>(nx)), m2(ny, std::vector<MHDStateVector>(nx));
        std::vector<std::vector<double> > phi(ny, std::vector<double>(nx));
        for (int i = 0; i < ny; i++)
            for (int j = 0; j < nx; j++)
            {
                m1[i][j] = mhdStateFrom(&in1[((size_t)i * nx + j) * 9]);
                m2[i][j] = mhdStateFrom(&in2[((size_t)i * nx + j) * 9]);
                phi[i][j] = phiIn[(size_t)i * nx + j];
            }
        MHDMultimaterialSystem initial(m1, m2, phi);
        MHDMaterialParameters q1 = p1, q2 = p2;
        double t = 0.0;
        int it = 0;
        while (t < tfinal && (maxSteps == 0 || it < maxSteps))
        {
            MHDMultimaterialSystem g = MHDRGFMSolver::applyRGFMBoundaryConditions2D(MHDMultimaterialSystem(m1, m2, phi), q1, q2);
            m1 = g.getMaterial1Cells2D();
            m2 = g.getMaterial2Cells2D();
            Grid b1 = MHDSolvers::insertBoundaryCells2D(m1, 2);
            Grid b2 = MHDSolvers::insertBoundaryCells2D(m2, 2);
            double dt = std::min(MHDSolvers::computeStableTimeStep2D(b1, dx, cfl, t, tfinal, it, q1),
                                 MHDSolvers::computeStableTimeStep2D(b2, dx, cfl, t, tfinal, it, q2));
            phi = MHDRGFMSolver::update2DLevelSetFunctionX(phi, dx, 0.5 * dt, m1, m2);
            phi = MHDRGFMSolver::update2DLevelSetFunctionY(phi, dx, dt, m1, m2);
            phi = MHDRGFMSolver::update2DLevelSetFunctionX(phi, dx, 0.5 * dt, m1, m2);
            q1.setHyperbolicWaveSpeed(4.0);
            q2.setHyperbolicWaveSpeed(4.0);
            q1.configureParabolicDamping();
            q2.configureParabolicDamping();
            Grid *mats[2] = {&m1, &m2};
            MHDMaterialParameters *qs[2] = {&q1, &q2};
            for (int k = 0; k < 2; k++)
            {
                Grid &m = *mats[k];
                for (int s = 0; s < sub; s++)
                {
                    Grid b = MHDSolvers::insertBoundaryCells2D(m, 1);
                    MHDForcingSolver::computeRungeKuttaTimeStep2D(m, b, dx, 0.5 * (dt / sub), bias, limiter, *qs[k]);
                }
                Grid b = MHDSolvers::insertBoundaryCells2D(m, 2);
                MHDSecondOrderSolver::computeXSLICTimeStep2D(m, b, dx, 0.5 * dt, bias, limiter, *qs[k]);
                b = MHDSolvers::insertBoundaryCells2D(m, 2);
                MHDSecondOrderSolver::computeYSLICTimeStep2D(m, b, dx, dt, bias, limiter, *qs[k]);
                b = MHDSolvers::insertBoundaryCells2D(m, 2);
                MHDSecondOrderSolver::computeXSLICTimeStep2D(m, b, dx, 0.5 * dt, bias, limiter, *qs[k]);
                for (int s = 0; s < sub; s++)
                {
                    Grid b2g = MHDSolvers::insertBoundaryCells2D(m, 1);
                    MHDForcingSolver::computeRungeKuttaTimeStep2D(m, b2g, dx, 0.5 * (dt / sub), bias, limiter, *qs[k]);
                }
            }
            t += dt;
            it += 1;
            dts.push_back(dt);
        }
        for (int i = 0; i < ny; i++)
            for (int j = 0; j < nx; j++)
            {
                putMhdState(m1[i][j], &out1[((size_t)i * nx + j) * 9]);
                putMhdState(m2[i][j], &out2[((size_t)i * nx + j) * 9]);
                outPhi[(size_t)i * nx + j] = phi[i][j];
            }
        if (checkApi && maxSteps == 0)
        {
            MHDMultimaterialSystem api = MHDRGFMSolver::solve2D(initial, dx, cfl, tfinal, bias, limiter, sub, 0, p1, p2);
            Grid a1 = api.getMaterial1Cells2D(), a2 = api.getMaterial2Cells2D();
            std::vector<std::vector<double> > ap = api.getLevelSetFunction2D();
            for (int i = 0; i < ny; i++)
                for (int j = 0; j < nx; j++)
                {
                    double r1[9], r2[9];
                    putMhdState(a1[i][j], r1);
                    putMhdState(a2[i][j], r2);
                    for (int f = 0; f < 9; f++)
                        match = match && sameBits(r1[f], out1[((size_t)i * nx + j) * 9 + f]) &&
                                sameBits(r2[f], out2[((size_t)i * nx + j) * 9 + f]);
                    match = match && sameBits(ap[i][j], outPhi[(size_t)i * nx + j]);
                }
        }
    }
    if (checkApi && maxSteps == 0) std::fprintf(stderr, "API_MATCH %d\n", match ? 1 : 0);
    std::fprintf(stderr, "STEPS %zu\n", dts.size());
    writeDoubles(argv[23], out1);
    writeDoubles(argv[24], out2);
    writeDoubles(argv[25], outPhi);
    writeDoubles(argv[26], dts);
    return 0;
}

// mhd_hllc_units n p1(4) p2(4) in out: per record of two states a, b:
// solveX(a, b) [9] | solveY(a, b) [9] | left, right, top, bottom star densities
// [4] | star x, y speeds [2].
int mhdHllcUnits(int argc, char **argv)
{
    if (argc < 13) { std::fprintf(stderr, "mhd_hllc_units: too few arguments\n"); return 2; }
    int n = std::atoi(argv[2]);
    MHDMaterialParameters p1 = mhdParams4(argv + 3);
    MHDMaterialParameters p2 = mhdParams4(argv + 7);
    std::vector<double> in = readDoubles(argv[11], (size_t)n * 18);
    std::vector<double> out((size_t)n * 24);
    for (int r = 0; r < n; r++)
    {
        MHDStateVector a = mhdStateFrom(&in[(size_t)r * 18]), b = mhdStateFrom(&in[(size_t)r * 18 + 9]);
        double *o = &out[(size_t)r * 24];
        putMhdState(MHDHLLCSolver::solveX(a, b, p1, p2), o);
        putMhdState(MHDHLLCSolver::solveY(a, b, p1, p2), o + 9);
        o[18] = MHDHLLCSolver::computeLeftStarRegionDensity(a, b, p1, p2);
        o[19] = MHDHLLCSolver::computeRightStarRegionDensity(a, b, p1, p2);
        o[20] = MHDHLLCSolver::computeTopStarRegionDensity(a, b, p1, p2);
        o[21] = MHDHLLCSolver::computeBottomStarRegionDensity(a, b, p1, p2);
        o[22] = MHDHLLCSolver::computeStarRegionXVelocity(a, b, p1, p2);
        o[23] = MHDHLLCSolver::computeStarRegionYVelocity(a, b, p1, p2);
    }
    writeDoubles(argv[12], out);
    return 0;
}

// The reference's own test drivers, end to end: each builds its initial
// condition, runs the reference's solve and writes its .dat files into the
// working directory (Euler/eulertests.cpp:7-52, RGFM/eulerrgfmtests.cpp,
// MHD/mhdtests.cpp:7-120, HPR/hprtests.cpp:7-57).
//   driver <name> <cellCount> <subcycles or reinitialisation frequency>
int driverRun(int argc, char **argv)
{
    if (argc < 5) { std::fprintf(stderr, "driver: too few arguments\n"); return 2; }
    const std::string name = argv[2];
    const int n = std::atoi(argv[3]), k = std::atoi(argv[4]);
    if (name == "euler_toro1") EulerTests::solveToroTest1(n);
    else if (name == "euler_2d_toro1") EulerTests::solve2DToroTest1(n);
    else if (name == "rgfm_toro1_hllc") EulerRGFMTests::solveToroTest1HLLC(n, k);
    else if (name == "rgfm_fedkiw_hllc") EulerRGFMTests::solveFedkiwTestHLLC(n, k);
    else if (name == "rgfm_2d_toro1_hllc") EulerRGFMTests::solve2DToroTest1HLLC(n, k);
    else if (name == "rgfm_2d_fedkiw_hllc") EulerRGFMTests::solve2DFedkiwTestHLLC(n, k);
    else if (name == "mhd_dumbser1") MHDTests::solveDumbserTest1(n, k);
    else if (name == "mhd_dumbser2") MHDTests::solveDumbserTest2(n, k);
    else if (name == "mhd_dumbser3") MHDTests::solveDumbserTest3(n, k);
    else if (name == "mhd_2d_dumbser1") MHDTests::solve2DDumbserTest1(n, k);
    else if (name == "mhd_2d_dumbser2") MHDTests::solve2DDumbserTest2(n, k);
    else if (name == "hpr_stokes") HPRTests::solveStokesFirstProblem(n, k);
    else if (name == "hpr_heat") HPRTests::solveHeatConductionProblem(n, k);
    else { std::fprintf(stderr, "driver: unknown %s\n", name.c_str()); return 2; }
    return 0;
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
    if (mode == "hpr_run_p") return hprRunP(argc, argv);
    if (mode == "hpr_units_p") return hprUnitsP(argc, argv);
    if (mode == "hpr_eos_units") return hprEosUnits(argc, argv);
    if (mode == "hpr_driver") return hprDriver(argc, argv);
    if (mode == "rgfm_run") return rgfmRun(argc, argv);
    if (mode == "hllc_units") return hllcUnits(argc, argv);
    if (mode == "dat") return datOut(argc, argv);
    if (mode == "driver") return driverRun(argc, argv);
    if (mode == "mp_run") return mpRun(argc, argv);
    if (mode == "mhd_rgfm_run") return mhdRgfmRun(argc, argv);
    if (mode == "mhd_hllc_units") return mhdHllcUnits(argc, argv);
    if (mode == "mp_units") return mpUnits(argc, argv);
    if (mode == "mp_reinit") return mpReinit(argc, argv);
    std::fprintf(stderr, "unknown mode %s\n", mode.c_str());
    return 2;
}
