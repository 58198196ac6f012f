// sg_medium.hip -- seeded random media on the GPU (spectral method).
//
// The reference builds its random velocity model with a spectral method
// (randomMedium/RandomMedia_3_d.m:20-86): amplitude sqrt(P(k)) with uniformly
// random phases (:47-57), an inverse 3-D FFT and its modulus (:65-66), the
// mean removed and the result scaled to a relative standard deviation (0.1,
// :74-85), V = v0 (1 + delta) (:87-93), written as raw float32 (V.dat,
// :101-103).  The same pipeline here, on one GPU:
//  * k_spectrum: one thread per frequency bin writes sqrt(P) e^{i theta};
//    theta comes from a counter-based generator (splitmix64 of the seed and
//    the bin index), so a seed gives the same medium on any device;
//  * hipFFT's single-precision C2C inverse transform, in place;
//  * k_moments: FP64 block sums of the field and its square (|psi| for the
//    script's own spectrum, Re psi for the Gaussian and von Karman spectra),
//    added on the host in block order: mean and standard deviation;
//  * k_finish: v0 (1 + eps (x - mean) / std) (or eps (x - mean) / std when
//    v0 <= 0) into float32, x fastest.
// k_spectrum evaluates the spectrum from per-axis tables (it is a function of
// three 1-D terms); the three kernels stream the complex array once each.
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/sgfd.h"
#include "sg_error.hpp"

namespace sg {
namespace {

constexpr int MTHREADS = 256;
constexpr int MBLOCKS = 8192;  // fixed grid of the reductions: sums do not depend on the device

__host__ __device__ inline uint64_t mix64(uint64_t x)
{
    x += 0x9e3779b97f4a7c15ULL;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
    return x ^ (x >> 31);
}

struct SpecArgs {
    int nx, ny, nz, kind;
    double abc, expo;  // REFERENCE: a0 a1 a2; VON_KARMAN: -(order + 3/2) / 2
    uint64_t s0;
};

// signed frequency index of bin m of an n-point transform (numpy.fft.fftfreq * n)
inline double sidx(int m, int n) { return (double)(m <= (n - 1) / 2 ? m : m - n); }

// Per-axis factors of the spectrum (host, FP64), so a point costs a few
// operations instead of three pow calls:
//   REFERENCE:  t = (k^2 a^2)^p, k = s / l       amp = sqrt(abc / (1 + tx + ty + tz))
//               (RandomMedia_3_d.m:29; the script squares b for z, with a = b = c)
//   GAUSSIAN:   t = exp(-k^2 a^2 / 8), k = 2 pi s / n  amp = tx ty tz = sqrt(exp(-q / 4))
//   VON_KARMAN: t = k^2 a^2, k = 2 pi s / n      amp = (1 + tx + ty + tz)^(-(order + 3/2) / 2)
void axis_table(const sg_spectrum &sp, int axis, int n, double *t)
{
    const double a = sp.a[axis];
    for (int m = 0; m < n; ++m) {
        const double s = sidx(m, n);
        if (sp.kind == SG_SPECTRUM_REFERENCE) {
            const double k = s / sp.l;
            t[m] = std::pow(k * k * a * a, sp.order);
        } else {
            const double k = 6.283185307179586 * s / n, q = k * k * a * a;
            t[m] = sp.kind == SG_SPECTRUM_GAUSSIAN ? std::exp(-0.125 * q) : q;
        }
    }
}

// One thread per bin, grid (x blocks, ny, nz) so no index is divided:
// sqrt(P) e^{i theta}, theta = 2 pi u with u the top 53 bits of
// mix64(s0 ^ bin) over 2^53, evaluated as sincospi(2u) in single precision
// (the stored spectrum is complex64).
__global__ void k_spectrum(SpecArgs a, const double *__restrict__ tab, float2 *__restrict__ psi)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
    if (i >= a.nx) return;
    const long long p = ((long long)k * a.ny + j) * a.nx + i;
    const double tx = tab[i], ty = tab[a.nx + j], tz = tab[a.nx + a.ny + k];
    double amp;
    if (a.kind == SG_SPECTRUM_REFERENCE) amp = sqrt(a.abc / (1.0 + tx + ty + tz));
    else if (a.kind == SG_SPECTRUM_GAUSSIAN) amp = tx * ty * tz;
    else amp = pow(1.0 + tx + ty + tz, a.expo);
    const double u2 = (double)(mix64(a.s0 ^ (uint64_t)p) >> 11) * (2.0 / 9007199254740992.0);
    float sn, cs;
    sincospif((float)u2, &sn, &cs);
    const float am = (float)amp;
    psi[p] = make_float2(am * cs, am * sn);
}

__device__ inline double field_value(float2 v, int take_abs)
{
    return take_abs ? sqrt((double)v.x * v.x + (double)v.y * v.y) : (double)v.x;
}

// per-block FP64 sums of x and x^2 (one pass; the mean is of the order of
// the standard deviation for both |psi| and Re psi, so E[x^2] - mean^2 keeps
// ~15 digits)
__global__ void k_moments(const float2 *__restrict__ psi, long long n, int take_abs, double *__restrict__ part)
{
    __shared__ double r1[MTHREADS], r2[MTHREADS];
    double s1 = 0, s2 = 0;
    for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < n; p += (long long)gridDim.x * blockDim.x) {
        const double x = field_value(psi[p], take_abs);
        s1 += x;
        s2 += x * x;
    }
    r1[threadIdx.x] = s1;
    r2[threadIdx.x] = s2;
    __syncthreads();
    for (int w = MTHREADS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            r1[threadIdx.x] += r1[threadIdx.x + w];
            r2[threadIdx.x] += r2[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = r1[0];
        part[2 * blockIdx.x + 1] = r2[0];
    }
}

__global__ void k_finish(const float2 *__restrict__ psi, long long n, int take_abs, double mean, double gain,
                         double v0, float *__restrict__ out)
{
    for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < n; p += (long long)gridDim.x * blockDim.x) {
        const double d = gain * (field_value(psi[p], take_abs) - mean);
        out[p] = (float)(v0 > 0 ? v0 * (1.0 + d) : d);
    }
}

}  // namespace
}  // namespace sg

using sg::fail;

namespace {

int random_medium(int nx, int ny, int nz, const sg_spectrum *spec, unsigned long long seed, double v0, double eps,
                  float *out, double *device_ms)
{
    if (!spec || !out || nx < 1 || ny < 1 || nz < 1 || !(eps > 0))
        return fail(SG_ERR_INVALID, "sg_random_medium: invalid arguments");
    const int kind = spec->kind;
    if (kind != SG_SPECTRUM_REFERENCE && kind != SG_SPECTRUM_GAUSSIAN && kind != SG_SPECTRUM_VON_KARMAN)
        return fail(SG_ERR_INVALID, "sg_random_medium: unknown spectrum kind");
    for (int a = 0; a < 3; ++a)
        if (!(spec->a[a] > 0)) return fail(SG_ERR_INVALID, "sg_random_medium: correlation lengths must be > 0");
    if (kind == SG_SPECTRUM_REFERENCE && !(spec->l > 0))
        return fail(SG_ERR_INVALID, "sg_random_medium: the reference spectrum needs l > 0");
    if (kind == SG_SPECTRUM_VON_KARMAN && !(spec->order > 0))
        return fail(SG_ERR_INVALID, "sg_random_medium: the von Karman order must be > 0");
    const long long n = (long long)nx * ny * nz;
    if (n < 2) return fail(SG_ERR_INVALID, "sg_random_medium: need at least two points");
    if (ny > 65535 || nz > 65535) return fail(SG_ERR_UNSUPPORTED, "sg_random_medium: ny and nz are limited to 65535");

    const sg::SpecArgs sa{nx, ny, nz, kind, spec->a[0] * spec->a[1] * spec->a[2], -0.5 * (spec->order + 1.5),
                          sg::mix64(seed)};
    const int take_abs = kind == SG_SPECTRUM_REFERENCE;
    std::vector<double> htab((size_t)nx + ny + nz);
    sg::axis_table(*spec, 0, nx, htab.data());
    sg::axis_table(*spec, 1, ny, htab.data() + nx);
    sg::axis_table(*spec, 2, nz, htab.data() + nx + ny);
    float2 *psi = nullptr;
    float *dout = nullptr;
    double *part = nullptr, *tab = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipfftHandle plan = 0;
    bool have_plan = false;
    auto cleanup = [&]() {
        if (have_plan) (void)hipfftDestroy(plan);
        if (psi) (void)hipFree(psi);
        if (dout) (void)hipFree(dout);
        if (part) (void)hipFree(part);
        if (tab) (void)hipFree(tab);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (st) (void)hipStreamDestroy(st);
    };
    auto hip_fail = [&](hipError_t e, const char *what) {
        cleanup();
        return fail(SG_ERR_HIP, std::string("sg_random_medium: ") + what + ": " + hipGetErrorString(e));
    };
    hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e != hipSuccess) return hip_fail(e, "stream");
    e = hipMalloc(&psi, (size_t)n * sizeof(float2));
    if (e == hipSuccess) e = hipMalloc(&dout, (size_t)n * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&part, sizeof(double) * 2 * sg::MBLOCKS);
    if (e == hipSuccess) e = hipMalloc(&tab, sizeof(double) * htab.size());
    if (e != hipSuccess) {
        cleanup();
        return fail(SG_ERR_NOMEM, "sg_random_medium: device allocation failed");
    }
    if ((e = hipMemcpy(tab, htab.data(), sizeof(double) * htab.size(), hipMemcpyHostToDevice)) != hipSuccess)
        return hip_fail(e, "table copy");
    // the plan (and any kernel set-up of the FFT library) before the timed
    // part; axes of one point are left out (a 3-D transform with unit axes is
    // the lower-rank transform of the others, in the same memory order)
    int dims[3], rank = 0;
    for (int d : {nz, ny, nx})
        if (d > 1) dims[rank++] = d;
    hipfftResult pr = rank == 3   ? hipfftPlan3d(&plan, dims[0], dims[1], dims[2], HIPFFT_C2C)
                      : rank == 2 ? hipfftPlan2d(&plan, dims[0], dims[1], HIPFFT_C2C)
                                  : hipfftPlan1d(&plan, dims[0], HIPFFT_C2C, 1);
    if (pr != HIPFFT_SUCCESS) {
        cleanup();
        return fail(SG_ERR_HIP, "sg_random_medium: FFT plan creation failed");
    }
    have_plan = true;
    if (hipfftSetStream(plan, st) != HIPFFT_SUCCESS) {
        cleanup();
        return fail(SG_ERR_HIP, "sg_random_medium: hipfftSetStream failed");
    }
    if ((e = hipEventRecord(e0, st)) != hipSuccess) return hip_fail(e, "event");
    const long long nb = (n + sg::MTHREADS - 1) / sg::MTHREADS;
    const int grid = (int)(nb < 65536 ? nb : 65536);
    hipLaunchKernelGGL(sg::k_spectrum, dim3((nx + sg::MTHREADS - 1) / sg::MTHREADS, ny, nz), dim3(sg::MTHREADS), 0, st,
                       sa, tab, psi);
    if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "k_spectrum");
    if (hipfftExecC2C(plan, (hipfftComplex *)psi, (hipfftComplex *)psi, HIPFFT_BACKWARD) != HIPFFT_SUCCESS) {
        cleanup();
        return fail(SG_ERR_HIP, "sg_random_medium: hipfftExecC2C failed");
    }

    // mean and population standard deviation (the script's sqrt(sum/N)),
    // block sums added in block order
    std::vector<double> h(2 * sg::MBLOCKS);
    hipLaunchKernelGGL(sg::k_moments, dim3(sg::MBLOCKS), dim3(sg::MTHREADS), 0, st, psi, n, take_abs, part);
    if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "k_moments");
    if ((e = hipMemcpyAsync(h.data(), part, sizeof(double) * h.size(), hipMemcpyDeviceToHost, st)) != hipSuccess ||
        (e = hipStreamSynchronize(st)) != hipSuccess)
        return hip_fail(e, "moments copy");
    double s1 = 0, s2 = 0;
    for (int b = 0; b < sg::MBLOCKS; ++b) {
        s1 += h[2 * b];
        s2 += h[2 * b + 1];
    }
    const double mean = s1 / (double)n, var = s2 / (double)n - mean * mean;
    const double mom[2] = {mean, var > 0 ? std::sqrt(var) : 0.0};
    // The complex64 field holds each value to ~6e-8 of the mean, so the
    // normalised field is accurate to ~6e-8 / (std / |mean|).  A spread
    // below 1e-4 of the mean (errors above ~6e-4 of eps) is refused rather
    // than scaled up to eps.
    if (!(mom[1] > 1e-4 * std::fabs(mom[0]))) {
        cleanup();
        return fail(SG_ERR_INVALID, "sg_random_medium: the field's spread is below 1e-4 of its mean, beyond single "
                                    "precision (spectrum too narrow for the grid)");
    }
    hipLaunchKernelGGL(sg::k_finish, dim3(grid), dim3(sg::MTHREADS), 0, st, psi, n, take_abs, mom[0], eps / mom[1],
                       v0, dout);
    if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "k_finish");
    if ((e = hipEventRecord(e1, st)) != hipSuccess) return hip_fail(e, "event");
    if ((e = hipMemcpyAsync(out, dout, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, st)) != hipSuccess ||
        (e = hipStreamSynchronize(st)) != hipSuccess)
        return hip_fail(e, "result copy");
    if (device_ms) {
        float ms = 0;
        if ((e = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess) return hip_fail(e, "event time");
        *device_ms = ms;
    }
    cleanup();
    return SG_OK;
}

}  // namespace

extern "C" int sg_random_medium(int nx, int ny, int nz, const sg_spectrum *spec, unsigned long long seed, double v0,
                                double eps, float *out)
{
    return random_medium(nx, ny, nz, spec, seed, v0, eps, out, nullptr);
}

extern "C" int sg_random_medium_timed(int nx, int ny, int nz, const sg_spectrum *spec, unsigned long long seed,
                                      double v0, double eps, float *out, double *device_ms)
{
    if (!device_ms) return fail(SG_ERR_INVALID, "sg_random_medium_timed: device_ms is NULL");
    return random_medium(nx, ny, nz, spec, seed, v0, eps, out, device_ms);
}
