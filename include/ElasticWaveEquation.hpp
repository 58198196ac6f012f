// ElasticWaveEquation.hpp -- header-only C++ drop-in for the reference driver
// class, built on the C ABI (sgfd.h).
//
// Mirrors 3DGPUwithPML_1DArraySingleGPU_concise/ElasticWaveEquation.h:17-60:
// the same constructor arguments (GRID is the TOTAL extent, PML and 5-cell
// halo included; Medium holds buoyancy/lambda/mu x-fastest; one station;
// DifferenceCoefficient; main frequency), run(), propagate(it, s),
// sourceFunction(it), stationDataOutput(), snapShootSlice(it) and
// snapShootVolum(it) (same files and text format).  GPUDim is accepted and
// ignored: the library picks its own launch shape.  The default source mode
// is the CUDA reference's (accumulating); pass SG_SOURCE_TRANSIENT for the
// CPU reference's behaviour.  Errors throw sg::Error instead of exiting.
#pragma once

#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "sgfd.h"

typedef struct GRID {
    int xLength, yLength, zLength;
} GRID;
typedef struct Delta {
    float t, x, y, z;
} Delta;
typedef struct SOURCELOCATION {
    int x, y, z;
} SOURCELOCATION;
typedef struct STATIONLOCATION {
    int x, y, z;
} STATIONLOCATION;
typedef struct Medium {
    float *buoyancy, *lambda, *mu;
} Medium;
typedef struct DifferenceCoefficient {
    float diff_coef[5];
} DifferenceCoefficient;
typedef struct GPUDim {
    unsigned blocksPerGrid[3], threadsPerBlock[3];
} GPUDim;

namespace sg {
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};
inline void check(int rc)
{
    if (rc != SG_OK) throw Error(rc, sg_last_error());
}
}  // namespace sg

class ElasticWaveEquation {
public:
    int timeLength;
    GRID grid;
    Delta delta;
    int pml;
    float mainFrequncy;
    SOURCELOCATION sourceLocationCPU;
    STATIONLOCATION stationLocationCPU;
    std::vector<float> stationDataCPU;

    ElasticWaveEquation(int timeLength, GPUDim /*gpuDim*/, int pml, GRID grid, Delta delta, Medium medCPU,
                        SOURCELOCATION sourceLocation, STATIONLOCATION stationLocation, DifferenceCoefficient dc,
                        float mainFrequncy, int source_mode = SG_SOURCE_ACCUMULATE, int precision = SG_FP32)
        : timeLength(timeLength), grid(grid), delta(delta), pml(pml), mainFrequncy(mainFrequncy),
          sourceLocationCPU(sourceLocation), stationLocationCPU(stationLocation)
    {
        sg_config c{};
        c.nx = grid.xLength;
        c.ny = grid.yLength;
        c.nz = grid.zLength;
        c.pml = pml;
        c.dt = delta.t;
        c.dx = delta.x;
        c.dy = delta.y;
        c.dz = delta.z;
        for (int i = 0; i < 5; ++i) c.diff_coef[i] = dc.diff_coef[i];
        c.time_length = timeLength;
        c.precision = precision;
        c.source_mode = source_mode;
        c.kernel = precision == SG_FP64 ? SG_KERNEL_SIMPLE : SG_KERNEL_TILED;  // FP64: per-point kernels
        sg::check(sg_create(&c, &s_));
        sg::check(sg_set_medium(s_, medCPU.buoyancy, medCPU.lambda, medCPU.mu));
        std::vector<float> w(timeLength);
        for (int it = 0; it < timeLength; ++it) w[it] = sourceFunction(it);
        sg::check(sg_set_source(s_, sourceLocation.x, sourceLocation.y, sourceLocation.z, w.data(), timeLength));
        const int st[3] = {stationLocation.x, stationLocation.y, stationLocation.z};
        sg::check(sg_set_stations(s_, st, 1, 1u << SG_SXX, 1));
    }
    ElasticWaveEquation(const ElasticWaveEquation &) = delete;
    ElasticWaveEquation &operator=(const ElasticWaveEquation &) = delete;
    ~ElasticWaveEquation() { sg_destroy(s_); }

    // the CUDA reference's damping: tau 4.3 and the local S velocity
    // sqrt(mu*b) of each cell (..._concise/ElasticMacro.h:21,26-27,
    // ElasticWaveEquation.cu:285-288); the default is the CPU reference's
    // per-axis profile
    void setCudaDamping(double tauCuda = 4.3)
    {
        std::vector<double> g[3];
        const int n3[3] = {grid.xLength, grid.yLength, grid.zLength};
        const float h3[3] = {delta.x, delta.y, delta.z};
        for (int a = 0; a < 3; ++a) {
            g[a].resize((size_t)n3[a]);
            sg::check(sg_damping_profile(n3[a], pml, h3[a], SG_DAMP_CUDA, 1.0, 1e-3, tauCuda, g[a].data()));
        }
        sg::check(sg_set_damping_local(s_, g[0].data(), g[1].data(), g[2].data()));
    }

    // Ricker of the reference (..._concise/ElasticWaveEquation.cu:187-192)
    float sourceFunction(int timeIndex)
    {
        std::vector<float> w(timeIndex + 1);
        sg::check(sg_ricker(mainFrequncy, delta.t, timeIndex + 1, SG_RICKER_CEIL_DELAY, w.data()));
        return w[timeIndex];
    }

    // one step with an explicit source value (..._concise/ElasticWaveEquation.cu:194-206)
    // The device injects the sample of its own step counter, so timeIndex
    // must be the next step: steps are taken in order, each once.
    void propagate(int timeIndex, float sourceValue)
    {
        if (timeIndex != sg_steps_done(s_))
            throw sg::Error(SG_ERR_STATE, "propagate: timeIndex " + std::to_string(timeIndex) +
                                              " is not the next step (" + std::to_string(sg_steps_done(s_)) + ")");
        sg::check(sg_set_source_sample(s_, timeIndex, sourceValue));
        sg::check(sg_step(s_, 1));
    }

    // all steps, slices through the source every snapEvery steps from step
    // snapFrom on (the reference: every 100 from 500), then the station
    // buffer to the host (..._concise/ElasticWaveEquation.cu:140-185).
    // snapEvery = 0: no snapshots.
    void run(int snapEvery = 100, int snapFrom = 500)
    {
        if (snapEvery <= 0) {
            sg::check(sg_run(s_));
        } else {
            for (int it = 0; it < timeLength; ++it) {
                sg::check(sg_step(s_, 1));
                if (it % snapEvery == 0 && it >= snapFrom) snapShootSlice(it);
            }
        }
        long long n = 0;
        sg::check(sg_get_stations(s_, nullptr, 0, &n));
        stationDataCPU.resize((size_t)n);
        sg::check(sg_get_stations(s_, stationDataCPU.data(), n, &n));
    }

    // vx, sxx and sxy on the three grid planes through the source, one value
    // per line, files velocityXSliceXX_YoZ_<it>.txt ... stressXYSliceXX_XoY_<it>.txt
    // (..._concise/ElasticWaveEquation.cu:929-1005)
    void snapShootSlice(int it, const std::string &dir = ".") const
    {
        const int nx = grid.xLength, ny = grid.yLength, nz = grid.zLength;
        const int sx = sourceLocationCPU.x, sy = sourceLocationCPU.y, sz = sourceLocationCPU.z;
        const int fields[3] = {SG_VX, SG_SXX, SG_SXY};
        const char *names[3] = {"velocityX", "stressXX", "stressXY"};
        std::vector<float> f((size_t)nx * ny * nz);
        for (int q = 0; q < 3; ++q) {
            sg::check(sg_get_field(s_, fields[q], f.data()));
            auto at = [&](int i, int j, int k) { return f[(size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k)]; };
            std::ofstream yoz(dir + "/" + names[q] + "SliceXX_YoZ_" + std::to_string(it) + ".txt");
            for (int k = 0; k < nz; ++k)
                for (int j = 0; j < ny; ++j) yoz << at(sx, j, k) << '\n';
            std::ofstream xoz(dir + "/" + names[q] + "SliceXX_XoZ_" + std::to_string(it) + ".txt");
            for (int k = 0; k < nz; ++k)
                for (int i = 0; i < nx; ++i) xoz << at(i, sy, k) << '\n';
            std::ofstream xoy(dir + "/" + names[q] + "SliceXX_XoY_" + std::to_string(it) + ".txt");
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) xoy << at(i, j, sz) << '\n';
        }
    }

    // sxx over the updated volume [5, n-5)^3, one value per line, file
    // snapshot_0_<it>.txt (..._concise/ElasticWaveEquation.cu:1007-1037)
    void snapShootVolum(int it, const std::string &dir = ".") const
    {
        const int nx = grid.xLength, ny = grid.yLength, nz = grid.zLength;
        std::vector<float> f((size_t)nx * ny * nz);
        sg::check(sg_get_field(s_, SG_SXX, f.data()));
        std::ofstream out(dir + "/snapshot_0_" + std::to_string(it) + ".txt");
        for (int k = 5; k < nz - 5; ++k)
            for (int j = 5; j < ny - 5; ++j)
                for (int i = 5; i < nx - 5; ++i) out << f[(size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k)] << '\n';
    }

    // binary variant (no counterpart in the reference): one field over the
    // whole grid as raw little-endian float32, x fastest, file
    // <field>_<it>.bin with field one of vx vy vz sxx syy szz sxy sxz syz
    void snapShootBinary(int it, int field, const std::string &dir = ".") const
    {
        static const char *names[9] = {"vx", "vy", "vz", "sxx", "syy", "szz", "sxy", "sxz", "syz"};
        if (field < 0 || field > 8) throw sg::Error(SG_ERR_RANGE, "snapShootBinary: bad field");
        std::vector<float> f((size_t)grid.xLength * grid.yLength * grid.zLength);
        sg::check(sg_get_field(s_, field, f.data()));
        std::ofstream out(dir + "/" + names[field] + "_" + std::to_string(it) + ".bin", std::ios::binary);
        out.write(reinterpret_cast<const char *>(f.data()), (std::streamsize)(f.size() * sizeof(float)));
    }

    // one value per line (..._concise/ElasticWaveEquation.cu:1040-1050)
    void stationDataOutput(const char *path = "stationData.txt") const
    {
        std::ofstream f(path);
        for (float v : stationDataCPU) f << v << '\n';
    }

    sg_solver *handle() { return s_; }

private:
    sg_solver *s_ = nullptr;
};
