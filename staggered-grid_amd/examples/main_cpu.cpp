// The reference's CPU main program (3DCPUwithPML_3DArray_GridIncludePMLandOrder/
// main.cpp:10-62) on the drop-in class of include/ElasticWaveEquation3D.hpp:
// interior 101^3, pml 10, 600 steps, density 2500, velocity Vs = 2000,
// lambda = rho*v^2, mu = rho*v^2/3, the source at the centre, the class
// defaults for spacing, time step and frequency.  It writes the reference's
// curveLine, snapshot and damping files into the working directory.
//
// Optional arguments: Lx Ly Lz pml T [sx sy sz [model_dir]] for other
// sizes; model_dir holds b.raw, lambda.raw and mu.raw, float32 over the total
// grid in [i][j][k] order (k fastest), instead of the homogeneous medium.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <ctime>
#include <string>
#include <vector>

#include "ElasticWaveEquation3D.hpp"

// a float*** view [i][j][k] over one contiguous block, k fastest
struct Array3D {
    std::vector<float> data;
    std::vector<float *> rows;
    std::vector<float **> planes;
    Array3D(int nx, int ny, int nz) : data((size_t)nx * ny * nz, 0.0f), rows((size_t)nx * ny), planes((size_t)nx)
    {
        for (int i = 0; i < nx; ++i) {
            for (int j = 0; j < ny; ++j) rows[(size_t)i * ny + j] = &data[((size_t)i * ny + j) * nz];
            planes[(size_t)i] = &rows[(size_t)i * ny];
        }
    }
    float ***ptr() { return planes.data(); }
};

static bool read_raw(const std::string &path, std::vector<float> &v)
{
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    const size_t got = std::fread(v.data(), sizeof(float), v.size(), f);
    std::fclose(f);
    return got == v.size();
}

int main(int argc, char **argv)
{
    GRID grid = {101, 101, 101};
    int pml = 10, timeLength = 600;
    if (argc >= 6) {
        grid = {std::atoi(argv[1]), std::atoi(argv[2]), std::atoi(argv[3])};
        pml = std::atoi(argv[4]);
        timeLength = std::atoi(argv[5]);
    }
    SOURCELOCATION sourceLocation = {grid.xLength / 2 + pml + 5, grid.yLength / 2 + pml + 5,
                                     grid.zLength / 2 + pml + 5};
    if (argc >= 9) sourceLocation = {std::atoi(argv[6]), std::atoi(argv[7]), std::atoi(argv[8])};
    const int nx = grid.xLength + 2 * pml + 10, ny = grid.yLength + 2 * pml + 10, nz = grid.zLength + 2 * pml + 10;

    Array3D buoyancy(nx, ny, nz), lambda(nx, ny, nz), mu(nx, ny, nz);
    if (argc >= 10) {
        const std::string d = argv[9];
        if (!read_raw(d + "/b.raw", buoyancy.data) || !read_raw(d + "/lambda.raw", lambda.data) ||
            !read_raw(d + "/mu.raw", mu.data)) {
            std::fprintf(stderr, "cannot read the model files in %s\n", d.c_str());
            return 1;
        }
    } else {
        const float density = 2500.0f, velocity = 2000.0f;
        for (size_t p = 0; p < buoyancy.data.size(); ++p) {
            buoyancy.data[p] = 1.0f / density;
            lambda.data[p] = density * std::pow(velocity, 2.0f);
            mu.data[p] = density * std::pow(velocity, 2.0f) / 3.0f;
        }
    }

    try {
        ElasticWaveEquation3D ewq(grid, timeLength, pml, buoyancy.ptr(), lambda.ptr(), mu.ptr(), sourceLocation);
        const time_t start = time(nullptr);
        ewq.propagate();
        const time_t stop = time(nullptr);
        std::printf("%ld\n", (long)(stop - start));
        std::printf("curve samples %zu\n", ewq.curveLineData[0].size());
    } catch (const sg::Error &e) {
        std::fprintf(stderr, "sgfd error %d: %s\n", e.code, e.what());
        return 1;
    }
    return 0;
}
