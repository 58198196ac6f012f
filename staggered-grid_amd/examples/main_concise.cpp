// The reference CUDA driver program (3DGPUwithPML_1DArraySingleGPU_concise/main.cu)
// ported to the drop-in header: same grid, medium, source, station and
// coefficients and damping; writes stationData.txt and the reference's slice
// snapshots.  Argument 1 (optional) = time steps.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "ElasticWaveEquation.hpp"

int main(int argc, char **argv)
{
    const int timeLength = argc > 1 ? std::atoi(argv[1]) : 1500;
    GRID grid = {300, 200, 100};
    Delta delta = {0.0005f, 10.0f, 10.0f, 10.0f};
    const int pml = 10;
    SOURCELOCATION sourceLocation = {grid.xLength / 2, grid.yLength / 2, grid.zLength / 2};
    STATIONLOCATION stationLocation = {grid.xLength / 2 + pml, grid.yLength / 2 + pml, grid.zLength / 2 + pml};
    DifferenceCoefficient dc = {{1.21124f, -8.97217e-2f, 1.38428e-2f, -1.76566e-3f, 1.18680e-4f}};
    const size_t n = (size_t)grid.xLength * grid.yLength * grid.zLength;
    std::vector<float> b(n, 1.0f / 2600.0f), lam(n, 2600.0f * 5000.0f * 5000.0f), mu(n, 2600.0f * 5000.0f * 5000.0f / 3.0f);
    Medium medCPU = {b.data(), lam.data(), mu.data()};
    try {
        ElasticWaveEquation ewq(timeLength, GPUDim{}, pml, grid, delta, medCPU, sourceLocation, stationLocation, dc,
                                20.0f);
        ewq.setCudaDamping();  // the CUDA reference's damping (tau 4.3, local S velocity)
        ewq.run();             // slices through the source every 100 steps from 500, as the reference
        ewq.stationDataOutput();
        std::printf("recorded %zu samples, sample[235] = %g\n", ewq.stationDataCPU.size(),
                    ewq.stationDataCPU.size() > 235 ? ewq.stationDataCPU[235] : 0.0f);
    } catch (const sg::Error &e) {
        std::fprintf(stderr, "sgfd error %d: %s\n", e.code, e.what());
        return 1;
    }
    return 0;
}
