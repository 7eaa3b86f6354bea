// vanguard_compat.hpp -- the reference's Euler solver classes over the C ABI.
//
// Same class names, static signatures and argument meaning as the reference's
// headers (Euler/eulermaterialparameters.h:5-37, Euler/eulerstatevector.h:8-62,
// Solvers/solvers.h:8-45, Solvers/secondordersolver.h:32-65,
// Solvers/firstordersolver.h:63-95), so a driver written against them -- the
// reference's own Euler test drivers, or an RGFM-style loop over the per-sweep
// calls -- recompiles unchanged and runs on the GPU.  Header-only; link with
// -lvanguard_hip.  Every solve runs in EXACT mode (the reference's arithmetic,
// bit for bit); failures throw std::runtime_error with vg_last_error().
//
// Only the state's primitive accessors are provided (the physics lives in the
// library); states convert to the C ABI's array of structures in the state
// vector's own field order (rho, u, v, w, p, lambda).
#ifndef VANGUARD_COMPAT_HPP
#define VANGUARD_COMPAT_HPP

#include <stdexcept>
#include <string>
#include <vector>

#include "vanguard.h"

using std::vector;

namespace vanguard_compat {

inline vg_ctx *context()
{
    thread_local struct Holder {
        vg_ctx *ctx = nullptr;
        ~Holder()
        {
            if (ctx) vg_ctx_destroy(ctx);
        }
    } h;
    if (!h.ctx && vg_ctx_create(&h.ctx, 0) != VG_OK) throw std::runtime_error(vg_last_error());
    return h.ctx;
}

inline void check(int rc)
{
    if (rc != VG_OK) throw std::runtime_error(std::string("vanguard: ") + vg_last_error());
}

}  // namespace vanguard_compat

class EulerMaterialParameters {
public:
    EulerMaterialParameters() : EulerMaterialParameters(1.0) {}
    EulerMaterialParameters(double newAdiabaticIndex) : EulerMaterialParameters(newAdiabaticIndex, 0.0) {}
    EulerMaterialParameters(double newAdiabaticIndex, double newStiffeningParameter)
        : EulerMaterialParameters(newAdiabaticIndex, newStiffeningParameter, 25.0, 0.0, 0.25, 250.0)
    {
    }
    EulerMaterialParameters(double newAdiabaticIndex, double newStiffeningParameter, double newSpecificHeatCapacity,
                            double newEnergyOfFormation, double newIgnitionTemperature, double newReactionRate)
        : p_{newAdiabaticIndex, newStiffeningParameter, newSpecificHeatCapacity, newEnergyOfFormation,
             newIgnitionTemperature, newReactionRate}
    {
    }

    void setAdiabaticIndex(double v) { p_.gamma = v; }
    void setStiffeningParameter(double v) { p_.p_inf = v; }
    void setSpecificHeatCapacity(double v) { p_.cv = v; }
    void setEnergyOfFormation(double v) { p_.e_form = v; }
    void setIgnitionTemperature(double v) { p_.t_ign = v; }
    void setReactionRate(double v) { p_.k_rate = v; }

    double getAdiabaticIndex() { return p_.gamma; }
    double getStiffeningParameter() { return p_.p_inf; }
    double getSpecificHeatCapacity() { return p_.cv; }
    double getEnergyOfFormation() { return p_.e_form; }
    double getIgnitionTemperature() { return p_.t_ign; }
    double getReactionRate() { return p_.k_rate; }

    const vg_euler_params *c_params() const { return &p_; }

private:
    vg_euler_params p_;
};

class EulerStateVector {
public:
    // The default state (1, 0, 0, 0, 1, 1) is the reference's 2D corner ghost.
    EulerStateVector() : EulerStateVector(1.0, 0.0, 0.0, 0.0, 1.0, 1.0) {}
    EulerStateVector(double newDensity, double newXVelocity, double newYVelocity, double newZVelocity,
                     double newPressure)
        : EulerStateVector(newDensity, newXVelocity, newYVelocity, newZVelocity, newPressure, 1.0)
    {
    }
    EulerStateVector(double newDensity, double newXVelocity, double newYVelocity, double newZVelocity,
                     double newPressure, double newReactionProgress)
        : w_{newDensity, newXVelocity, newYVelocity, newZVelocity, newPressure, newReactionProgress}
    {
    }

    void setPrimitiveVariableVector(vector<double> v)
    {
        for (int k = 0; k < 6; k++) w_[k] = v[k];
    }
    vector<double> computePrimitiveVariableVector() { return vector<double>(w_, w_ + 6); }

    void setDensity(double v) { w_[0] = v; }
    void setXVelocity(double v) { w_[1] = v; }
    void setYVelocity(double v) { w_[2] = v; }
    void setZVelocity(double v) { w_[3] = v; }
    void setPressure(double v) { w_[4] = v; }
    void setReactionProgress(double v) { w_[5] = v; }

    double getDensity() { return w_[0]; }
    double getXVelocity() { return w_[1]; }
    double getYVelocity() { return w_[2]; }
    double getZVelocity() { return w_[3]; }
    double getPressure() { return w_[4]; }
    double getReactionProgress() { return w_[5]; }

    const double *data() const { return w_; }
    double *data() { return w_; }

private:
    double w_[6];
};

namespace vanguard_compat {

inline vector<double> pack(const vector<EulerStateVector> &c)
{
    vector<double> a(c.size() * 6);
    for (size_t i = 0; i < c.size(); i++)
        for (int k = 0; k < 6; k++) a[i * 6 + k] = c[i].data()[k];
    return a;
}
inline vector<double> pack2d(const vector<vector<EulerStateVector> > &c)
{
    const size_t ny = c.size(), nx = ny ? c[0].size() : 0;
    vector<double> a(ny * nx * 6);
    for (size_t i = 0; i < ny; i++)
        for (size_t j = 0; j < nx; j++)
            for (int k = 0; k < 6; k++) a[(i * nx + j) * 6 + k] = c[i][j].data()[k];
    return a;
}
inline void unpack(const vector<double> &a, vector<EulerStateVector> &c)
{
    for (size_t i = 0; i < c.size(); i++)
        for (int k = 0; k < 6; k++) c[i].data()[k] = a[i * 6 + k];
}
inline void unpack2d(const vector<double> &a, vector<vector<EulerStateVector> > &c)
{
    const size_t ny = c.size(), nx = ny ? c[0].size() : 0;
    for (size_t i = 0; i < ny; i++)
        for (size_t j = 0; j < nx; j++)
            for (int k = 0; k < 6; k++) c[i][j].data()[k] = a[(i * nx + j) * 6 + k];
}

// The maximum of max(|u|, |v|) + c (2D) or |u| + c (1D) over exactly the cells
// given, computed on the device (periodic index map: no extra ghost values).
inline double max_wave_speed(int dim, int ny, int nx, const vector<double> &a, EulerMaterialParameters &m)
{
    vg_euler_state *st = nullptr;
    check(vg_euler_state_create(context(), dim, ny, nx, &st));
    int rc = vg_euler_state_load(st, a.data(), 0);
    double s = 0.0;
    if (rc == VG_OK) rc = vg_euler_state_max_wave_speed(st, VG_BC_PERIODIC, m.c_params(), &s);
    vg_euler_state_destroy(st);
    check(rc);
    return s;
}

}  // namespace vanguard_compat

class Solvers {
public:
    // Solvers::insertBoundaryCells (solvers.cpp:7-32): one layer copied, two
    // layers mirror-ordered.
    static vector<EulerStateVector> insertBoundaryCells(vector<EulerStateVector> &currentCells, int boundarySize)
    {
        const int n = (int)currentCells.size();
        vector<EulerStateVector> out(n + 2 * boundarySize);
        if (boundarySize == 1) {
            out[0] = currentCells[0];
            out[n + 1] = currentCells[n - 1];
        } else if (boundarySize == 2) {
            out[0] = currentCells[1];
            out[1] = currentCells[0];
            out[n + 2] = currentCells[n - 1];
            out[n + 3] = currentCells[n - 2];
        }
        for (int i = 0; i < n; i++) out[i + boundarySize] = currentCells[i];
        return out;
    }

    // Solvers::insertBoundaryCells2D (solvers.cpp:88-143): corners keep the default state.
    static vector<vector<EulerStateVector> > insertBoundaryCells2D(vector<vector<EulerStateVector> > &currentCells,
                                                                   int boundarySize)
    {
        const int ny = (int)currentCells.size(), nx = (int)currentCells[0].size(), b = boundarySize;
        vector<vector<EulerStateVector> > out(ny + 2 * b, vector<EulerStateVector>(nx + 2 * b));
        if (b == 1) {
            for (int i = 0; i < ny; i++) {
                out[i + 1][0] = currentCells[i][0];
                out[i + 1][nx + 1] = currentCells[i][nx - 1];
            }
            for (int j = 0; j < nx; j++) {
                out[0][j + 1] = currentCells[0][j];
                out[ny + 1][j + 1] = currentCells[ny - 1][j];
            }
        } else if (b == 2) {
            for (int i = 0; i < ny; i++) {
                out[i + 2][0] = currentCells[i][1];
                out[i + 2][1] = currentCells[i][0];
                out[i + 2][nx + 2] = currentCells[i][nx - 1];
                out[i + 2][nx + 3] = currentCells[i][nx - 2];
            }
            for (int j = 0; j < nx; j++) {
                out[0][j + 2] = currentCells[1][j];
                out[1][j + 2] = currentCells[0][j];
                out[ny + 2][j + 2] = currentCells[ny - 1][j];
                out[ny + 3][j + 2] = currentCells[ny - 2][j];
            }
        }
        for (int i = 0; i < ny; i++)
            for (int j = 0; j < nx; j++) out[i + b][j + b] = currentCells[i][j];
        return out;
    }

    // computeMaximumWaveSpeed[2D] (solvers.cpp:259-276, 318-339) over the cells given.
    static double computeMaximumWaveSpeed(vector<EulerStateVector> &currentCells,
                                          EulerMaterialParameters materialParameters)
    {
        return vanguard_compat::max_wave_speed(1, 1, (int)currentCells.size(), vanguard_compat::pack(currentCells),
                                               materialParameters);
    }
    static double computeMaximumWaveSpeed2D(vector<vector<EulerStateVector> > &currentCells,
                                            EulerMaterialParameters materialParameters)
    {
        return vanguard_compat::max_wave_speed(2, (int)currentCells.size(), (int)currentCells[0].size(),
                                               vanguard_compat::pack2d(currentCells), materialParameters);
    }

    // computeStableTimeStep (solvers.cpp:389-395, 413-419, 437-452).
    static double computeStableTimeStep(double timeStep, double currentTime, double finalTime, int currentIteration)
    {
        double newTimeStep = timeStep;
        if (currentIteration <= 5) newTimeStep *= 0.2;
        if ((currentTime + newTimeStep) > finalTime) newTimeStep = finalTime - currentTime;
        return newTimeStep;
    }
    static double computeStableTimeStep(vector<EulerStateVector> &currentCells, double cellSpacing,
                                        double CFLCoefficient, double currentTime, double finalTime,
                                        int currentIteration, EulerMaterialParameters materialParameters)
    {
        const double s = computeMaximumWaveSpeed(currentCells, materialParameters);
        return computeStableTimeStep(CFLCoefficient * (cellSpacing / s), currentTime, finalTime, currentIteration);
    }
    static double computeStableTimeStep2D(vector<vector<EulerStateVector> > &currentCells, double cellSpacing,
                                          double CFLCoefficient, double currentTime, double finalTime,
                                          int currentIteration, EulerMaterialParameters materialParameters)
    {
        const double s = computeMaximumWaveSpeed2D(currentCells, materialParameters);
        return computeStableTimeStep(CFLCoefficient * (cellSpacing / s), currentTime, finalTime, currentIteration);
    }
};

namespace vanguard_compat {

// A solve whose final state holds NaN or inadmissible cells still returns it,
// as the reference does (NaN propagates silently, secondordersolver.cpp:295);
// any other failure throws.
inline void check_solve(int rc)
{
    if (rc != VG_ERR_NONFINITE) check(rc);
}

inline vector<EulerStateVector> solve1d(vector<EulerStateVector> &initialCells, double cellSpacing, double cfl,
                                        double finalTime, double bias, int limiter, int order,
                                        EulerMaterialParameters &m)
{
    vector<double> in = pack(initialCells), out(in.size());
    vg_run_stats st;
    check_solve(vg_euler_solve_1d(context(), in.data(), out.data(), (int)initialCells.size(), cellSpacing, cfl,
                                  finalTime, 0, bias, limiter, order, VG_BC_TRANSMISSIVE, m.c_params(),
                                  VG_MODE_EXACT, nullptr, 0, &st));
    vector<EulerStateVector> cells(initialCells.size());
    unpack(out, cells);
    return cells;
}

inline vector<vector<EulerStateVector> > solve2d(vector<vector<EulerStateVector> > &initialCells, double cellSpacing,
                                                 double cfl, double finalTime, double bias, int limiter, int order,
                                                 EulerMaterialParameters &m)
{
    const int ny = (int)initialCells.size(), nx = (int)initialCells[0].size();
    vector<double> in = pack2d(initialCells), out(in.size());
    vg_run_stats st;
    check_solve(vg_euler_solve_2d(context(), in.data(), out.data(), ny, nx, cellSpacing, cfl, finalTime, 0, bias,
                                  limiter, order, VG_BC_TRANSMISSIVE, m.c_params(), VG_MODE_EXACT, nullptr, 0, &st));
    vector<vector<EulerStateVector> > cells(ny, vector<EulerStateVector>(nx));
    unpack2d(out, cells);
    return cells;
}

// One per-sweep call: currentCells is overwritten from the ghosted snapshot.
inline void time_step_1d(vector<EulerStateVector> &currentCells, vector<EulerStateVector> &withBoundary,
                         double cellSpacing, double timeStep, double bias, int limiter, int order,
                         EulerMaterialParameters &m)
{
    vector<double> wb = pack(withBoundary), out(currentCells.size() * 6);
    check(vg_euler_time_step_1d(context(), out.data(), wb.data(), (int)currentCells.size(), cellSpacing, timeStep,
                                bias, limiter, order, m.c_params(), VG_MODE_EXACT));
    unpack(out, currentCells);
}

inline void time_step_2d(vector<vector<EulerStateVector> > &currentCells,
                         vector<vector<EulerStateVector> > &withBoundary, int dir, double cellSpacing, double timeStep,
                         double bias, int limiter, int order, EulerMaterialParameters &m)
{
    const int ny = (int)currentCells.size(), nx = (int)currentCells[0].size();
    vector<double> wb = pack2d(withBoundary), out((size_t)ny * nx * 6);
    check(vg_euler_time_step_2d(context(), out.data(), wb.data(), ny, nx, dir, cellSpacing, timeStep, bias, limiter,
                                order, m.c_params(), VG_MODE_EXACT));
    unpack2d(out, currentCells);
}

}  // namespace vanguard_compat

// SLIC (Solvers/secondordersolver.h:32-65).  subcyclingIterations has no effect
// for Euler (an empty loop in the reference, secondordersolver.cpp:302-305).
class SecondOrderSolver {
public:
    static void computeSLICTimeStep(vector<EulerStateVector> &currentCells,
                                    vector<EulerStateVector> &currentCellsWithBoundary, double cellSpacing,
                                    double timeStep, double bias, int slopeLimiter,
                                    EulerMaterialParameters materialParameters)
    {
        vanguard_compat::time_step_1d(currentCells, currentCellsWithBoundary, cellSpacing, timeStep, bias, slopeLimiter,
                                      2, materialParameters);
    }
    static void computeXSLICTimeStep2D(vector<vector<EulerStateVector> > &currentCells,
                                       vector<vector<EulerStateVector> > &currentCellsWithBoundary, double cellSpacing,
                                       double timeStep, double bias, int slopeLimiter,
                                       EulerMaterialParameters materialParameters)
    {
        vanguard_compat::time_step_2d(currentCells, currentCellsWithBoundary, 0, cellSpacing, timeStep, bias,
                                      slopeLimiter, 2, materialParameters);
    }
    static void computeYSLICTimeStep2D(vector<vector<EulerStateVector> > &currentCells,
                                       vector<vector<EulerStateVector> > &currentCellsWithBoundary, double cellSpacing,
                                       double timeStep, double bias, int slopeLimiter,
                                       EulerMaterialParameters materialParameters)
    {
        vanguard_compat::time_step_2d(currentCells, currentCellsWithBoundary, 1, cellSpacing, timeStep, bias,
                                      slopeLimiter, 2, materialParameters);
    }
    static vector<EulerStateVector> solve(vector<EulerStateVector> &initialCells, double cellSpacing,
                                          double CFLCoefficient, double finalTime, double bias, int slopeLimiter,
                                          int subcyclingIterations, EulerMaterialParameters materialParameters)
    {
        (void)subcyclingIterations;
        return vanguard_compat::solve1d(initialCells, cellSpacing, CFLCoefficient, finalTime, bias, slopeLimiter, 2,
                                        materialParameters);
    }
    static vector<vector<EulerStateVector> > solve2D(vector<vector<EulerStateVector> > &initialCells,
                                                     double cellSpacing, double CFLCoefficient, double finalTime,
                                                     double bias, int slopeLimiter, int subcyclingIterations,
                                                     EulerMaterialParameters materialParameters)
    {
        (void)subcyclingIterations;
        return vanguard_compat::solve2d(initialCells, cellSpacing, CFLCoefficient, finalTime, bias, slopeLimiter, 2,
                                        materialParameters);
    }
};

// FORCE (Solvers/firstordersolver.h:63-95).
class FirstOrderSolver {
public:
    static void computeFORCETimeStep(vector<EulerStateVector> &currentCells,
                                     vector<EulerStateVector> &currentCellsWithBoundary, double cellSpacing,
                                     double timeStep, EulerMaterialParameters materialParameters)
    {
        vanguard_compat::time_step_1d(currentCells, currentCellsWithBoundary, cellSpacing, timeStep, 0.0, 0, 1,
                                      materialParameters);
    }
    static void computeXFORCETimeStep2D(vector<vector<EulerStateVector> > &currentCells,
                                        vector<vector<EulerStateVector> > &currentCellsWithBoundary,
                                        double cellSpacing, double timeStep, EulerMaterialParameters materialParameters)
    {
        vanguard_compat::time_step_2d(currentCells, currentCellsWithBoundary, 0, cellSpacing, timeStep, 0.0, 0, 1,
                                      materialParameters);
    }
    static void computeYFORCETimeStep2D(vector<vector<EulerStateVector> > &currentCells,
                                        vector<vector<EulerStateVector> > &currentCellsWithBoundary,
                                        double cellSpacing, double timeStep, EulerMaterialParameters materialParameters)
    {
        vanguard_compat::time_step_2d(currentCells, currentCellsWithBoundary, 1, cellSpacing, timeStep, 0.0, 0, 1,
                                      materialParameters);
    }
    static vector<EulerStateVector> solve(vector<EulerStateVector> &initialCells, double cellSpacing,
                                          double CFLCoefficient, double finalTime, int subcyclingIterations,
                                          EulerMaterialParameters materialParameters)
    {
        (void)subcyclingIterations;
        return vanguard_compat::solve1d(initialCells, cellSpacing, CFLCoefficient, finalTime, 0.0, 0, 1,
                                        materialParameters);
    }
    static vector<vector<EulerStateVector> > solve2D(vector<vector<EulerStateVector> > &initialCells,
                                                     double cellSpacing, double CFLCoefficient, double finalTime,
                                                     int subcyclingIterations,
                                                     EulerMaterialParameters materialParameters)
    {
        (void)subcyclingIterations;
        return vanguard_compat::solve2d(initialCells, cellSpacing, CFLCoefficient, finalTime, 0.0, 0, 1,
                                        materialParameters);
    }
};

#endif  // VANGUARD_COMPAT_HPP
