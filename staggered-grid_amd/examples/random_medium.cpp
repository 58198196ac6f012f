// The reference's random-model workflow through the C ABI and the drop-in
// header: build a velocity model with the spectral method of
// randomMedium/RandomMedia_3_d.m (its spectrum and parameters, seeded, on the
// GPU), write it as V.dat the way the script does, read it back as a solver
// would, turn it into a Medium (Vs = V / sqrt(3), rho 2000) and run the
// reference driver class on it.  Argument 1: output directory.
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "ElasticWaveEquation.hpp"

int main(int argc, char **argv)
{
    const std::string dir = argc > 1 ? argv[1] : ".";
    const int nx = 96, ny = 96, nz = 96;
    const size_t n = (size_t)nx * ny * nz;
    const std::string vdat = dir + "/V.dat";
    std::vector<float> v(n);
    sg_spectrum sp = {SG_SPECTRUM_REFERENCE, {10.0, 10.0, 10.0}, 10.0, 50.0};
    if (sg_random_medium(nx, ny, nz, &sp, 7ull, 2000.0, 0.1, v.data()) != SG_OK ||
        sg_write_raw_f32(vdat.c_str(), (long long)n, v.data()) != SG_OK ||
        sg_read_raw_f32(vdat.c_str(), (long long)n, v.data()) != SG_OK) {
        std::fprintf(stderr, "sgfd error: %s\n", sg_last_error());
        return 1;
    }
    double mean = 0, var = 0;
    for (float x : v) mean += x;
    mean /= (double)n;
    for (float x : v) var += (x - mean) * (x - mean);
    std::printf("V.dat: mean %.3f std %.3f\n", mean, std::sqrt(var / (double)n));

    const float rho = 2000.0f;
    std::vector<float> b(n, 1.0f / rho), lam(n), mu(n);
    for (size_t i = 0; i < n; ++i) {
        const float vs = v[i] / std::sqrt(3.0f);
        mu[i] = rho * vs * vs;
        lam[i] = rho * v[i] * v[i] - 2.0f * mu[i];
    }
    GRID grid = {nx, ny, nz};
    Delta delta = {0.001f, 10.0f, 10.0f, 10.0f};
    DifferenceCoefficient dc = {{1.21124f, -8.97217e-2f, 1.38428e-2f, -1.76566e-3f, 1.18680e-4f}};
    Medium med = {b.data(), lam.data(), mu.data()};
    try {
        ElasticWaveEquation ewq(200, GPUDim{}, 10, grid, delta, med, SOURCELOCATION{48, 48, 48},
                                STATIONLOCATION{60, 48, 48}, dc, 15.0f);
        ewq.run(0);  // all steps, no snapshots
        ewq.stationDataOutput((dir + "/stationData.txt").c_str());
    } catch (const sg::Error &e) {
        std::fprintf(stderr, "sgfd error %d: %s\n", e.code, e.what());
        return 1;
    }
    std::printf("run done\n");
    return 0;
}
