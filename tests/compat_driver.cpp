// A driver in the reference's style, compiled unchanged against the C++
// compatibility header (include/vanguard_compat.hpp) and run on the GPU:
// Toro's test 1 as Euler/eulertests.cpp:7-27 builds it (here with the 1000-cell
// minbee, CFL 0.9 configuration), solved (1) with SecondOrderSolver::solve and
// (2) with the solve loop written out over the per-sweep calls, as the
// reference's own solve body and RGFM solver do (secondordersolver.cpp:288-314,
// rgfmsolver.cpp:757-781).  Writes both final states as raw float64 into the
// directory given.
#include <cstdio>
#include <string>

#include "vanguard_compat.hpp"

static void write_cells(const std::string &path, vector<EulerStateVector> &cells)
{
    FILE *f = std::fopen(path.c_str(), "wb");
    for (size_t i = 0; i < cells.size(); i++) {
        double w[6] = {cells[i].getDensity(), cells[i].getXVelocity(), cells[i].getYVelocity(),
                       cells[i].getZVelocity(), cells[i].getPressure(), cells[i].getReactionProgress()};
        std::fwrite(w, sizeof(double), 6, f);
    }
    std::fclose(f);
}

int main(int argc, char **argv)
{
    if (argc != 2) return 2;
    const std::string dir = argv[1];
    int cellCount = 1000;
    double cellSpacing = 1.0 / cellCount;

    vector<EulerStateVector> initialCells(cellCount);
    EulerMaterialParameters materialParameters(1.4);

    for (int i = 0; i < cellCount; i++) {
        if (i <= (0.5 * cellCount)) {
            initialCells[i] = EulerStateVector(1.0, 0.0, 0.0, 0.0, 1.0);
        } else {
            initialCells[i] = EulerStateVector(0.125, 0.0, 0.0, 0.0, 0.1);
        }
    }

    vector<EulerStateVector> solution = SecondOrderSolver::solve(initialCells, cellSpacing, 0.9, 0.25, 0.0, 2, 0,
                                                                 materialParameters);
    write_cells(dir + "/compat_solve.f64", solution);

    double currentTime = 0.0;
    int currentIteration = 0;
    vector<EulerStateVector> currentCells = initialCells;
    while (currentTime < 0.25) {
        vector<EulerStateVector> currentCellsWithBoundary = Solvers::insertBoundaryCells(currentCells, 2);
        double timeStep = Solvers::computeStableTimeStep(currentCellsWithBoundary, cellSpacing, 0.9, currentTime, 0.25,
                                                         currentIteration, materialParameters);
        SecondOrderSolver::computeSLICTimeStep(currentCells, currentCellsWithBoundary, cellSpacing, timeStep, 0.0, 2,
                                               materialParameters);
        currentTime += timeStep;
        currentIteration += 1;
    }
    write_cells(dir + "/compat_loop.f64", currentCells);
    std::printf("compat driver: %d steps, t = %.17g\n", currentIteration, currentTime);
    return 0;
}
