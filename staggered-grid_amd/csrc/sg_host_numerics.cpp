// sg_host_numerics.cpp -- host-side FP64 numerics of the C ABI: staggered
// coefficients, damping profiles and the Ricker wavelet.  No device work.
#include <cmath>
#include <cstring>
#include <string>

#include "../../include/sgfd.h"
#include "sg_error.hpp"

// The wavelet must round exactly like the reference host code (no FMA).
#pragma clang fp contract(off)

extern "C" int sg_diff_coef(int kind, float out[5])
{
    if (!out) return sg::fail(SG_ERR_INVALID, "sg_diff_coef: out is NULL");
    if (kind == 0) {
        // DIFF_10 literals (3DCPUwithPML_3DArray_GridIncludePMLandOrder/ElasticMacro.h:65-69),
        // also DifferenceCoefficient in 3DGPUwithPML_1DArraySingleGPU_concise/main.cu:105-109
        const float c[5] = {1.21124f, -8.97217e-2f, 1.38428e-2f, -1.76566e-3f, 1.18680e-4f};
        std::memcpy(out, c, sizeof c);
    } else if (kind == 1) {
        // closed form of the 10th-order staggered first derivative
        out[0] = (float)(19845.0 / 16384.0);
        out[1] = (float)(-735.0 / 8192.0);
        out[2] = (float)(567.0 / 40960.0);
        out[3] = (float)(-405.0 / 229376.0);
        out[4] = (float)(35.0 / 294912.0);
    } else {
        return sg::fail(SG_ERR_INVALID, "sg_diff_coef: kind must be 0 or 1");
    }
    return SG_OK;
}

extern "C" int sg_damping_profile(int n, int pml, double h, int kind, double v, double r, double tau, double *out)
{
    if (!out || n < 11 || pml < 0 || 2 * pml > n - 10 || !(h > 0) || !(r > 0 && r < 1))
        return sg::fail(SG_ERR_INVALID, "sg_damping_profile: invalid arguments");
    // Distance into the layer as in calculateDamping
    // (3DCPUwithPML.../ElasticWaveEquation3D.cpp:271-296): pml at the outermost
    // updated cell, 1 next to the interior; zero elsewhere.
    double d0;
    if (kind == SG_DAMP_REFERENCE)
        d0 = std::log10(1.0 / r) * tau * v / (pml * h);  // ElasticMacro.h:21
    else if (kind == SG_DAMP_QUADRATIC)
        d0 = -3.0 * v * std::log(r) / (2.0 * pml * h);
    else if (kind == SG_DAMP_CUDA)  // ..._concise/ElasticMacro.h:26
        d0 = tau * v / h * (8.0 / 15.0 - 0.03 * pml + pml * (double)pml / 1500.0);
    else
        return sg::fail(SG_ERR_INVALID, "sg_damping_profile: unknown kind");
    for (int i = 0; i < n; ++i) {
        int dist = 0;
        if (i >= 5 && i < 5 + pml)
            dist = 5 + pml - i;
        else if (i >= n - 5 - pml && i < n - 5)
            dist = i - (n - 5 - pml) + 1;
        const double q = pml > 0 ? (double)dist / pml : 0.0;
        out[i] = dist ? d0 * q * q : 0.0;
    }
    return SG_OK;
}

extern "C" int sg_ricker(float f0, float dt, int nt, int delay_mode, float *out)
{
    if (!out || nt < 0 || !(f0 > 0) || !(dt > 0)) return sg::fail(SG_ERR_INVALID, "sg_ricker: invalid arguments");
    // sourceFunction (3DCPUwithPML.../ElasticWaveEquation3D.cpp:114-119) in its
    // own float arithmetic: t0 = ceil(1/(f0 dt)); tmp = pow(PI f0 ((it-t0) dt - 1/f0), 2)
    // evaluated in double and stored float; s = (1 - 2 tmp) exp(-tmp) in float.
    const float PI = 3.1415926f;
    const float t0 = delay_mode == SG_RICKER_CEIL_DELAY ? std::ceil(1.0f / (f0 * dt)) : 0.0f;
    for (int it = 0; it < nt; ++it) {
        const float arg = PI * f0 * (((float)it - t0) * dt - 1 / f0);
        const float tmp = (float)std::pow((double)arg, 2.0);
        out[it] = (1 - 2 * tmp) * std::exp(-tmp);
    }
    return SG_OK;
}

// ---------------------------------------------------------------------------
// Seeded Gaussian-correlated random field (host, multi-threaded, deterministic
// for any thread count).  White noise from a counter-based generator
// (splitmix64 of the point index, Box-Muller), filtered separably along x, y
// and z with g(r) = exp(-2 r^2 / a^2), |r| <= ceil(2a), periodic wrap, so the
// field's correlation is exp(-r^2/a^2); each pass is divided by sqrt(sum g^2)
// so the variance is 1; finally clipped to +-clip.  The reference's generator
// (randomMedium/RandomMedia_3_d.m) is an unseeded MATLAB spectral method that
// no solver reads; this one is seeded so runs are reproducible.
// ---------------------------------------------------------------------------
#include <thread>
#include <vector>
#include <cstdint>

namespace {

inline uint64_t splitmix64(uint64_t x)
{
    x += 0x9e3779b97f4a7c15ULL;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
    return x ^ (x >> 31);
}

template <class F>
void parallel_for(long long n, F fn)
{
    unsigned nt = std::thread::hardware_concurrency();
    if (nt == 0) nt = 1;
    if (nt > 16) nt = 16;
    if (n < 4096) nt = 1;
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t)
        th.emplace_back([=]() {
            for (long long i = (long long)t * n / nt; i < (long long)(t + 1) * n / nt; ++i) fn(i);
        });
    for (auto &x : th) x.join();
}

}  // namespace

extern "C" int sg_gaussian_field(int nx, int ny, int nz, double corr, unsigned long long seed, double clip, float *out)
{
    if (!out || nx < 1 || ny < 1 || nz < 1 || !(corr > 0) || !(clip > 0))
        return sg::fail(SG_ERR_INVALID, "sg_gaussian_field: invalid arguments");
    const long long N = (long long)nx * ny * nz;
    const uint64_t s0 = splitmix64(seed);
    parallel_for((N + 1) / 2, [&](long long p) {
        const uint64_t a = splitmix64(s0 ^ (uint64_t)(2 * p)), b = splitmix64(s0 ^ (uint64_t)(2 * p + 1));
        const double u1 = ((a >> 11) + 1.0) * (1.0 / 9007199254740993.0);  // (0, 1]
        const double u2 = (b >> 11) * (1.0 / 9007199254740992.0);
        const double r = std::sqrt(-2.0 * std::log(u1)), t = 6.283185307179586 * u2;
        out[2 * p] = (float)(r * std::cos(t));
        if (2 * p + 1 < N) out[2 * p + 1] = (float)(r * std::sin(t));
    });
    const int R = (int)std::ceil(2.0 * corr);
    std::vector<double> g(2 * R + 1);
    double s2 = 0;
    for (int q = -R; q <= R; ++q) {
        g[q + R] = std::exp(-2.0 * q * q / (corr * corr));
        s2 += g[q + R] * g[q + R];
    }
    for (auto &v : g) v /= std::sqrt(s2);
    const int n3[3] = {nx, ny, nz};
    const long long st3[3] = {1, nx, (long long)nx * ny};
    for (int ax = 0; ax < 3; ++ax) {
        const int n = n3[ax];
        const long long st = st3[ax];
        const long long lines = N / n;
        parallel_for(lines, [&](long long L) {
            // line L: base index with the filtered axis at 0
            long long base;
            if (ax == 0) base = L * nx;
            else if (ax == 1) base = (L / nx) * ((long long)nx * ny) + (L % nx);
            else base = L;
            std::vector<double> in(n);
            for (int i = 0; i < n; ++i) in[i] = out[base + i * st];
            for (int i = 0; i < n; ++i) {
                double acc = 0;
                for (int q = -R; q <= R; ++q) {
                    int m = (i + q) % n;
                    if (m < 0) m += n;
                    acc += g[q + R] * in[m];
                }
                out[base + i * st] = (float)acc;
            }
        });
    }
    parallel_for(N, [&](long long i) {
        const float c = (float)clip;
        out[i] = out[i] > c ? c : (out[i] < -c ? -c : out[i]);
    });
    return SG_OK;
}

// ---------------------------------------------------------------------------
// Raw float32 model files (V.dat of randomMedium/RandomMedia_3_d.m:101-103):
// n native little-endian floats, no header.
// ---------------------------------------------------------------------------
#include <cstdio>

extern "C" int sg_elastic_from_velocity(long long n, const float *vp, const float *vs, const float *rho,
                                        float *b, float *lam, float *mu)
{
    if (n < 0 || !vp || !vs || !rho || !b || !lam || !mu)
        return sg::fail(SG_ERR_INVALID, "sg_elastic_from_velocity: invalid arguments");
    for (long long i = 0; i < n; ++i) {
        if (!(rho[i] > 0) || !std::isfinite(rho[i]) || !(vs[i] > 0) || !std::isfinite(vs[i]) || !std::isfinite(vp[i]))
            return sg::fail(SG_ERR_INVALID, "sg_elastic_from_velocity: rho and vs must be positive and finite, vp "
                                            "finite (first bad index " + std::to_string(i) + ")");
        const double r = rho[i], s = vs[i], p = vp[i];
        const double m = r * s * s;
        b[i] = (float)(1.0 / r);
        mu[i] = (float)m;
        lam[i] = (float)(r * p * p - 2.0 * m);
    }
    return SG_OK;
}

extern "C" int sg_read_raw_f32(const char *path, long long n, float *out)
{
    if (!path || !out || n < 1) return sg::fail(SG_ERR_INVALID, "sg_read_raw_f32: invalid arguments");
    FILE *f = std::fopen(path, "rb");
    if (!f) return sg::fail(SG_ERR_INVALID, std::string("sg_read_raw_f32: cannot open ") + path);
    long long size = -1;
    if (std::fseek(f, 0, SEEK_END) == 0) size = (long long)std::ftell(f);
    if (size != n * (long long)sizeof(float)) {
        std::fclose(f);
        return sg::fail(SG_ERR_INVALID, std::string("sg_read_raw_f32: ") + path + " holds " + std::to_string(size) +
                                            " bytes, expected " + std::to_string(n * (long long)sizeof(float)));
    }
    std::rewind(f);
    const size_t got = std::fread(out, sizeof(float), (size_t)n, f);
    std::fclose(f);
    if (got != (size_t)n) return sg::fail(SG_ERR_INVALID, std::string("sg_read_raw_f32: short read of ") + path);
    return SG_OK;
}

extern "C" int sg_write_raw_f32(const char *path, long long n, const float *in)
{
    if (!path || !in || n < 1) return sg::fail(SG_ERR_INVALID, "sg_write_raw_f32: invalid arguments");
    FILE *f = std::fopen(path, "wb");
    if (!f) return sg::fail(SG_ERR_INVALID, std::string("sg_write_raw_f32: cannot create ") + path);
    const size_t put = std::fwrite(in, sizeof(float), (size_t)n, f);
    const bool ok = std::fclose(f) == 0 && put == (size_t)n;
    return ok ? SG_OK : sg::fail(SG_ERR_INVALID, std::string("sg_write_raw_f32: write to ") + path + " failed");
}

// Largest normalised Courant number of a model: the staggered operator's
// largest symbol is 2/h * sum |c_n| (at k h = pi, where sin((2n-1) k h / 2)
// alternates in sign like c_n), so the leapfrog scheme is stable for
// dt * max(vp) * sum |c_n| * sqrt(1/dx^2 + 1/dy^2 + 1/dz^2) <= 1, with
// vp = sqrt((lambda + 2 mu) b) per point.
extern "C" int sg_courant_number(long long n, const float *b, const float *lam, const float *mu, float dt, float dx,
                                 float dy, float dz, const float diff_coef[5], double *courant)
{
    if (n < 1 || !b || !lam || !mu || !diff_coef || !courant || !(dt > 0) || !(dx > 0) || !(dy > 0) || !(dz > 0))
        return sg::fail(SG_ERR_INVALID, "sg_courant_number: invalid arguments");
    double vp2 = 0.0;
    for (long long i = 0; i < n; ++i) {
        if (!(b[i] > 0) || !std::isfinite(b[i]) || !std::isfinite(lam[i]) || !(mu[i] > 0) || !std::isfinite(mu[i]) ||
            !(3.0 * (double)lam[i] + 2.0 * (double)mu[i] > 0))
            return sg::fail(SG_ERR_INVALID, "buoyancy and mu must be positive and finite, lambda finite with "
                                            "3 lambda + 2 mu > 0 (first bad index " + std::to_string(i) + ")");
        const double v = ((double)lam[i] + 2.0 * (double)mu[i]) * (double)b[i];
        if (v > vp2) vp2 = v;
    }
    double cs = 0.0;
    for (int k = 0; k < 5; ++k) cs += std::fabs((double)diff_coef[k]);
    const double hx = dx, hy = dy, hz = dz;
    *courant = (double)dt * std::sqrt(vp2) * cs * std::sqrt(1.0 / (hx * hx) + 1.0 / (hy * hy) + 1.0 / (hz * hz));
    return SG_OK;
}
