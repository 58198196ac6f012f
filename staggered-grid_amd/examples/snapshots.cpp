// Snapshot output through the drop-in header: a small heterogeneous-free
// model, slices through the source every 20 steps from step 20 (as the
// reference's run() does every 100 from 500), then the sxx volume and a
// binary vz volume at the last step.  Argument 1: output directory.
#include <cstdio>
#include <string>
#include <vector>

#include "ElasticWaveEquation.hpp"

int main(int argc, char **argv)
{
    const std::string dir = argc > 1 ? argv[1] : ".";
    const int timeLength = 41;
    GRID grid = {80, 70, 60};
    Delta delta = {0.0005f, 10.0f, 10.0f, 10.0f};
    const int pml = 8;
    SOURCELOCATION src = {40, 35, 30};
    STATIONLOCATION st = {50, 40, 35};
    DifferenceCoefficient dc = {{1.21124f, -8.97217e-2f, 1.38428e-2f, -1.76566e-3f, 1.18680e-4f}};
    const size_t n = (size_t)grid.xLength * grid.yLength * grid.zLength;
    std::vector<float> b(n, 1.0f / 2600.0f), lam(n, 2600.0f * 5000.0f * 5000.0f), mu(n, 2600.0f * 5000.0f * 5000.0f / 3.0f);
    Medium med = {b.data(), lam.data(), mu.data()};
    try {
        ElasticWaveEquation ewq(timeLength, GPUDim{}, pml, grid, delta, med, src, st, dc, 20.0f);
        for (int it = 0; it < timeLength; ++it) {
            ewq.propagate(it, ewq.sourceFunction(it));
            if (it % 20 == 0 && it >= 20) ewq.snapShootSlice(it, dir);
        }
        ewq.snapShootVolum(timeLength - 1, dir);
        ewq.snapShootBinary(timeLength - 1, SG_VZ, dir);
        std::printf("snapshots written\n");
    } catch (const sg::Error &e) {
        std::fprintf(stderr, "sgfd error %d: %s\n", e.code, e.what());
        return 1;
    }
    return 0;
}
