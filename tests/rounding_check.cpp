// Host check of the stochastic-rounding helpers of the 16-bit storage path
// (staggered-grid_amd/csrc/sg_precision.hpp): every result is one of the two
// representable neighbours of x, exact values and non-finite values are kept,
// and the mean over many random draws equals x (unbiased) to within the
// sampling error.  Prints "ok" or the first violation.
#include <cmath>
#include <cstdio>
#include <cstring>

#include "sg_precision.hpp"

static float h2f(_Float16 h) { return (float)h; }

static int check_half(float x)
{
    const int n = 200000;
    double sum = 0.0;
    const float lo = h2f((_Float16)x);
    float a = lo, b = lo;
    for (int k = 0; k < n; ++k) {
        const float v = h2f(sg::half_sr(x, sg::hash32(0x1234567u + 7919u * (unsigned)k)));
        if (v != lo) {
            if (b == lo && a == lo) b = v;
            else if (v != b) {
                std::printf("half %.9g: third value %.9g (%.9g %.9g)\n", x, v, a, b);
                return 1;
            }
        }
        sum += v;
    }
    const double mean = sum / n, gap = std::fabs((double)b - (double)a);
    if ((double)x < std::fmin(a, b) || (double)x > std::fmax(a, b)) {
        std::printf("half %.9g outside its neighbours %.9g %.9g\n", x, a, b);
        return 1;
    }
    // sampling error of a two-point variable: gap * 0.5 / sqrt(n); allow 5x
    if (std::fabs(mean - x) > 5.0 * 0.5 * gap / std::sqrt((double)n) + 1e-30) {
        std::printf("half %.9g biased: mean %.9g (gap %.3g)\n", x, mean, gap);
        return 1;
    }
    return 0;
}

static int check_bf16(float x)
{
    const int n = 200000;
    double sum = 0.0;
    for (int k = 0; k < n; ++k) {
        const uint32_t u = (uint32_t)sg::bf16_sr_bits(x, sg::hash32(0xabcdefu + 104729u * (unsigned)k)) << 16;
        float v;
        std::memcpy(&v, &u, 4);
        sum += v;
    }
    uint32_t ux;
    std::memcpy(&ux, &x, 4);
    const uint32_t t = ux & 0xffff0000u;
    float lo, hi;
    std::memcpy(&lo, &t, 4);
    const uint32_t t2 = t + 0x10000u;
    std::memcpy(&hi, &t2, 4);
    const double gap = std::fabs((double)hi - (double)lo);
    if (std::fabs(sum / n - x) > 5.0 * 0.5 * gap / std::sqrt((double)n) + 1e-30) {
        std::printf("bf16 %.9g biased: mean %.9g\n", x, sum / n);
        return 1;
    }
    return 0;
}

int main()
{
    const float xs[] = {1.0f, 1.0001f, 0.3333333f, -2.7182818f, 1000.37f, 65000.3f, -0.000123457f,
                        6.0e-5f, 1.0e-6f, -3.0e-8f, 2.0e-8f, 1.0e-9f, 7.77e-3f, -1234.5678f};
    for (float x : xs)
        if (check_half(x) || check_bf16(x)) return 1;
    // exact and special values are returned unchanged
    const float ex[] = {0.0f, -0.0f, 1.5f, -65504.0f, 0.00006103515625f};
    for (float x : ex)
        for (unsigned r : {0u, 0x80000000u, 0xffffffffu})
            if (h2f(sg::half_sr(x, r)) != x || std::signbit(h2f(sg::half_sr(x, r))) != std::signbit(x)) {
                std::printf("exact %.9g changed\n", x);
                return 1;
            }
    if (!std::isinf(h2f(sg::half_sr(1.0e6f, 0xffffffffu))) || !std::isnan(h2f(sg::half_sr(NAN, 123u)))) {
        std::printf("non-finite handling\n");
        return 1;
    }
    std::printf("ok\n");
    return 0;
}
