// ElasticWaveEquation3D.hpp -- header-only C++ drop-in for the reference's CPU
// driver class (3DCPUwithPML_3DArray_GridIncludePMLandOrder/
// ElasticWaveEquation3D.h:11-52, ElasticWaveEquation3D.cpp:12-311), built on
// the C ABI (sgfd.h).
//
// The CPU convention, as in the reference:
//  * GRID is the INTERIOR extent.  Every array spans L + 2*pml + 10 points per
//    axis (the PML and a 5-cell halo on each side, main.cpp:26-30).
//  * buoyancy, lambda and mu are float*** indexed [i][j][k] with k fastest
//    (ElasticMacro.h:116-127), owned by the caller.  They are read when
//    propagate() starts; the reference reads them during propagate(), so
//    changes made before that call are seen by both.
//  * The source location is in total-grid indices (main.cpp:15-17).
//  * The constructor defaults are the reference's: 5 m spacing, 0.5 ms, 30 Hz.
//  * The numerics are the reference's: the DIFF_10 coefficient literals, the
//    CPU damping formula with tau 3.4, Vs 2000 and R 0.001 (ElasticMacro.h:
//    16-23, .cpp:253-300), and the Ricker with the ceil(1/(f0*dt)) delay
//    (.cpp:114-119).  The source adds pow(dt,2)*s(it) to sxx, syy and szz
//    before the velocity update (.cpp:121-130).  The stress update rebuilds
//    sigma from its splits, so each injection reaches one velocity update
//    (the transient source).
//
// propagate() writes the reference's files into `dir`:
//  * curveLine1..3.txt: sxx at (pml+10, sy, sz), (sx, pml+10, sz) and
//    (sx, sy, pml+10) after every step, one value per line (.cpp:55-83).
//  * snapshot_0_<it+1>.txt: sxx over the whole grid, i outermost and k
//    innermost, after step it when it+1 >= 200 and it % 50 == 0, and after
//    step 299 (.cpp:71-80, 87-112).  snapshot_1..5 are created empty, as in
//    the reference.
//  * dampingX/Y/Z.txt: the three damping volumes (.cpp:302-311).
// writeFiles = false skips the files (the curve data stays in curveLineData).
// echo = true prints the reference's console lines: "Time Iterator: it", the
// source value, and the snapshot file names.  propagate() runs once per
// object: a second call would restart the reference's time index on live
// fields, which the library's step counter does not do, so it throws.
// Errors throw sg::Error instead of exiting.
#pragma once

#include <iostream>
#include <string>
#include <vector>

#include "ElasticWaveEquation.hpp"

class ElasticWaveEquation3D {
public:
    bool writeFiles = true;
    bool echo = false;
    std::string dir = ".";
    std::vector<float> curveLineData[3];  // the three curves after propagate()

    ElasticWaveEquation3D(GRID grid, int timeLength, int pml, float ***buoyancy, float ***lambda, float ***mu,
                          SOURCELOCATION sourceLocation, float deltaX = 5.0f, float deltaY = 5.0f,
                          float deltaZ = 5.0f, float deltaT = 0.0005f, float mainFrequncy = 30.0f,
                          int precision = SG_FP32)
        : timeLength(timeLength), deltaX(deltaX), deltaY(deltaY), deltaZ(deltaZ), deltaT(deltaT),
          mainFrequncy(mainFrequncy), grid(grid), pml(pml), sourceLocation(sourceLocation), buoyancy(buoyancy),
          lambda(lambda), mu(mu)
    {
        if (timeLength < 0) throw sg::Error(SG_ERR_INVALID, "ElasticWaveEquation3D: timeLength must be >= 0");
        if (!buoyancy || !lambda || !mu) throw sg::Error(SG_ERR_INVALID, "ElasticWaveEquation3D: NULL model array");
        n_[0] = grid.xLength + 2 * pml + 10;
        n_[1] = grid.yLength + 2 * pml + 10;
        n_[2] = grid.zLength + 2 * pml + 10;
        sg_config c{};
        c.nx = n_[0];
        c.ny = n_[1];
        c.nz = n_[2];
        c.pml = pml;
        c.dt = deltaT;
        c.dx = deltaX;
        c.dy = deltaY;
        c.dz = deltaZ;
        sg::check(sg_diff_coef(0, c.diff_coef));  // the DIFF_10 literals
        c.time_length = timeLength;
        c.precision = precision;
        c.source_mode = SG_SOURCE_TRANSIENT;
        c.kernel = precision == SG_FP64 ? SG_KERNEL_SIMPLE : SG_KERNEL_TILED;  // FP64: per-point kernels
        sg::check(sg_create(&c, &s_));
        try {
            // R, tau and Vs are float constants in the reference
            const double h3[3] = {deltaX, deltaY, deltaZ};
            for (int a = 0; a < 3; ++a) {
                damp_[a].resize((size_t)n_[a]);
                sg::check(sg_damping_profile(n_[a], pml, h3[a], SG_DAMP_REFERENCE, (double)2000.0f, (double)0.001f,
                                             (double)3.4f, damp_[a].data()));
            }
            sg::check(sg_set_damping(s_, damp_[0].data(), damp_[1].data(), damp_[2].data()));
            if (timeLength > 0) {
                std::vector<float> w((size_t)timeLength);
                sg::check(sg_ricker(mainFrequncy, deltaT, timeLength, SG_RICKER_CEIL_DELAY, w.data()));
                sg::check(sg_set_source(s_, sourceLocation.x, sourceLocation.y, sourceLocation.z, w.data(),
                                        timeLength));
            }
            const int sx = sourceLocation.x, sy = sourceLocation.y, sz = sourceLocation.z, e = pml + 10;
            const int st[9] = {e, sy, sz, sx, e, sz, sx, sy, e};
            sg::check(sg_set_stations(s_, st, 3, 1u << SG_SXX, 1));
        } catch (...) {
            sg_destroy(s_);
            throw;
        }
    }
    ElasticWaveEquation3D(const ElasticWaveEquation3D &) = delete;
    ElasticWaveEquation3D &operator=(const ElasticWaveEquation3D &) = delete;
    ~ElasticWaveEquation3D() { sg_destroy(s_); }

    // the time loop (.cpp:53-85): damping files, then per step the source,
    // the velocity and stress updates, the three curve samples and the
    // snapshots; the curve files at the end
    void propagate()
    {
        if (propagated_) throw sg::Error(SG_ERR_STATE, "ElasticWaveEquation3D: propagate() runs once per object");
        propagated_ = true;
        load_model();
        calculateDamping();
        int done = 0;
        for (int it = 0; it < timeLength; ++it) {
            const bool snap = ((it + 1) >= 200 && it % 50 == 0) || (it + 1) == 300;
            if (echo) std::cout << "Time Iterator: " << it << '\n' << sourceFunction(it) << std::endl;
            if (echo || snap || it + 1 == timeLength) {
                sg::check(sg_step(s_, it + 1 - done));
                done = it + 1;
            }
            if (snap) snapshoot(it + 1);
        }
        long long n = 0;
        sg::check(sg_get_stations(s_, nullptr, 0, &n));
        std::vector<float> st((size_t)n);
        if (n > 0) sg::check(sg_get_stations(s_, st.data(), n, &n));
        for (int c = 0; c < 3; ++c) {
            curveLineData[c].resize((size_t)(n / 3));
            for (long long r = 0; r < n / 3; ++r) curveLineData[c][(size_t)r] = st[(size_t)(3 * r + c)];
            if (writeFiles) {
                std::ofstream f(path("curveLine" + std::to_string(c + 1) + ".txt"));
                for (float v : curveLineData[c]) f << v << '\n';
            }
        }
    }

    // sxx over the whole grid, i outermost, k innermost (.cpp:87-112)
    void snapshoot(int it) const
    {
        if (echo)
            for (int q = 0; q < 6; ++q) std::cout << "snapshot_" << q << "_" << it << ".txt" << std::endl;
        if (!writeFiles) return;
        std::vector<float> f((size_t)n_[0] * n_[1] * n_[2]);
        sg::check(sg_get_field(s_, SG_SXX, f.data()));
        for (int q = 1; q < 6; ++q) {
            std::ofstream empty(path("snapshot_" + std::to_string(q) + "_" + std::to_string(it) + ".txt"));
        }
        std::ofstream out(path("snapshot_0_" + std::to_string(it) + ".txt"));
        for (int i = 0; i < n_[0]; ++i)
            for (int j = 0; j < n_[1]; ++j)
                for (int k = 0; k < n_[2]; ++k) out << f[at(i, j, k)] << '\n';
    }

    // Ricker of the reference (.cpp:114-119)
    float sourceFunction(int timeIndex) const
    {
        if (timeIndex < 0) throw sg::Error(SG_ERR_RANGE, "sourceFunction: negative time index");
        std::vector<float> w((size_t)timeIndex + 1);
        sg::check(sg_ricker(mainFrequncy, deltaT, timeIndex + 1, SG_RICKER_CEIL_DELAY, w.data()));
        return w[(size_t)timeIndex];
    }

    // the three damping volumes, one value per line, i outermost
    // (.cpp:253-311): damping_a(i,j,k) = d_a(index on axis a) inside the
    // updated box [5, n-5)^3, zero on the halo (d_a is zero off the PML)
    void calculateDamping() const
    {
        if (!writeFiles) return;
        static const char *names[3] = {"dampingX.txt", "dampingY.txt", "dampingZ.txt"};
        for (int a = 0; a < 3; ++a) {
            std::ofstream out(path(names[a]));
            for (int i = 0; i < n_[0]; ++i)
                for (int j = 0; j < n_[1]; ++j)
                    for (int k = 0; k < n_[2]; ++k) {
                        const int ijk[3] = {i, j, k};
                        const bool in = updated(0, i) && updated(1, j) && updated(2, k);
                        out << (in ? (float)damp_[a][(size_t)ijk[a]] : 0.0f) << '\n';
                    }
        }
    }

    sg_solver *handle() { return s_; }

private:
    int timeLength;
    float deltaX, deltaY, deltaZ, deltaT;
    float mainFrequncy;
    GRID grid;
    int pml;
    SOURCELOCATION sourceLocation;
    float ***buoyancy, ***lambda, ***mu;
    sg_solver *s_ = nullptr;
    int n_[3] = {0, 0, 0};
    std::vector<double> damp_[3];
    bool propagated_ = false;

    size_t at(int i, int j, int k) const { return (size_t)i + (size_t)n_[0] * ((size_t)j + (size_t)n_[1] * k); }
    bool updated(int a, int i) const { return i >= 5 && i < n_[a] - 5; }
    std::string path(const std::string &name) const { return dir + "/" + name; }

    // [i][j][k] (k fastest) -> the library's x-fastest layout
    void load_model()
    {
        const size_t n = (size_t)n_[0] * n_[1] * n_[2];
        std::vector<float> b(n), lam(n), m(n);
        for (int i = 0; i < n_[0]; ++i)
            for (int j = 0; j < n_[1]; ++j)
                for (int k = 0; k < n_[2]; ++k) {
                    const size_t p = at(i, j, k);
                    b[p] = buoyancy[i][j][k];
                    lam[p] = lambda[i][j][k];
                    m[p] = mu[i][j][k];
                }
        sg::check(sg_set_medium(s_, b.data(), lam.data(), m.data()));
    }
};
