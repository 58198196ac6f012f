// sg_precision.hpp -- storage types and their load/store conversions.
//
// FP32/FP64 store the value itself.  FP16/BF16 store value*scale with a
// per-array power-of-two scale (so scaling is exact) and convert to float for
// all arithmetic; stores round to nearest even.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sg {

struct bf16 {
    uint16_t bits;
};

template <class T, class C>
struct Cvt;

template <>
struct Cvt<float, float> {
    static __device__ __forceinline__ float ld(float x, float) { return x; }
    static __device__ __forceinline__ float st(float v, float) { return v; }
};

template <>
struct Cvt<double, double> {
    static __device__ __forceinline__ double ld(double x, float) { return x; }
    static __device__ __forceinline__ double st(double v, float) { return v; }
};

template <>
struct Cvt<_Float16, float> {
    static __device__ __forceinline__ float ld(_Float16 x, float inv) { return (float)x * inv; }
    static __device__ __forceinline__ _Float16 st(float v, float scale) { return (_Float16)(v * scale); }
};

template <>
struct Cvt<bf16, float> {
    static __device__ __forceinline__ float ld(bf16 x, float inv)
    {
        return __builtin_bit_cast(float, (uint32_t)x.bits << 16) * inv;
    }
    static __device__ __forceinline__ bf16 st(float v, float scale)
    {
        // round to nearest even on the upper 16 bits; NaN stays NaN
        uint32_t u = __builtin_bit_cast(uint32_t, v * scale);
        if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return bf16{(uint16_t)((u >> 16) | 0x40)};
        u += 0x7fffu + ((u >> 16) & 1u);
        return bf16{(uint16_t)(u >> 16)};
    }
};

// 32-bit integer hash (the fmix32 finalizer of MurmurHash3): the counter-based
// random source of stochastic rounding, keyed by (point, step, field) so that
// results do not depend on launch shape or thread order.
__host__ __device__ __forceinline__ uint32_t hash32(uint32_t h)
{
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}

__host__ __device__ __forceinline__ uint32_t round_key(uint32_t point, uint32_t step, uint32_t field)
{
    return hash32(point * 0x9e3779b9u ^ hash32(step * 16u + field));
}

// 16 random bits for storing field f at (point, step): fields 2m and 2m+1
// take the low and the high half of one hash, so a point costs 6 hashes
// instead of 9 (the stochastic-rounding overhead was mostly hashing).
__host__ __device__ __forceinline__ uint32_t sr_bits(uint32_t point, uint32_t step, uint32_t field)
{
    const uint32_t k = round_key(point, step, field >> 1);
    return (field & 1u) ? k >> 16 : k & 0xffffu;
}

// Stochastic rounding of a float to FP16: the nearer neighbour x_n (round to
// nearest even) or the farther one x_f, with P(x_f) = |x - x_n| / |x_f - x_n|
// (resolved to 2^-16), so the stored value is unbiased.  r: random bits, the
// low 16 used.
__host__ __device__ __forceinline__ _Float16 half_sr(float x, uint32_t r)
{
    const _Float16 h = (_Float16)x;
    const float xh = (float)h;
    const float d = x - xh;
    if (d == 0.0f || !(xh <= 65504.0f && xh >= -65504.0f)) return h;
    const uint16_t hb = __builtin_bit_cast(uint16_t, h);
    uint16_t fb;
    if ((hb & 0x7fffu) == 0) fb = (uint16_t)(d > 0.0f ? 0x0001u : 0x8001u);  // from zero: smallest subnormal
    else fb = (uint16_t)(((d > 0.0f) == (xh > 0.0f)) ? hb + 1 : hb - 1);     // one step away from / towards zero
    const _Float16 f = __builtin_bit_cast(_Float16, fb);
    const float p = d / ((float)f - xh);                                        // in (0, 1/2]
    return ((float)(r & 0xffffu) * (1.0f / 65536.0f) < p) ? f : h;
}

// Stochastic rounding to BF16: add 16 random bits below the kept mantissa
// and truncate (finite values).
__host__ __device__ __forceinline__ uint16_t bf16_sr_bits(float x, uint32_t r)
{
    uint32_t u = __builtin_bit_cast(uint32_t, x);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0x007fffffu) ? 0x40u : 0u));
    u += r & 0xffffu;
    return (uint16_t)(u >> 16);
}

template <class T, class C>
struct CvtSR {
    static __device__ __forceinline__ T st(C v, float, uint32_t) { return (T)v; }
};

template <>
struct CvtSR<_Float16, float> {
    static __device__ __forceinline__ _Float16 st(float v, float scale, uint32_t r) { return half_sr(v * scale, r); }
};

template <>
struct CvtSR<bf16, float> {
    static __device__ __forceinline__ bf16 st(float v, float scale, uint32_t r) { return bf16{bf16_sr_bits(v * scale, r)}; }
};

// explicit fused multiply-add: the tiled sweeps are compiled with contraction
// off so every instantiation (tile shape, interior/PML) rounds identically
__device__ __forceinline__ float fmad(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmad(double a, double b, double c) { return __builtin_fma(a, b, c); }

}  // namespace sg
