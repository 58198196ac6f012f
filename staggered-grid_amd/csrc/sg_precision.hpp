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
        return __uint_as_float((uint32_t)x.bits << 16) * inv;
    }
    static __device__ __forceinline__ bf16 st(float v, float scale)
    {
        // round to nearest even on the upper 16 bits; NaN stays NaN
        uint32_t u = __float_as_uint(v * scale);
        if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return bf16{(uint16_t)((u >> 16) | 0x40)};
        u += 0x7fffu + ((u >> 16) & 1u);
        return bf16{(uint16_t)(u >> 16)};
    }
};

// explicit fused multiply-add: the tiled sweeps are compiled with contraction
// off so every instantiation (tile shape, interior/PML) rounds identically
__device__ __forceinline__ float fmad(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmad(double a, double b, double c) { return __builtin_fma(a, b, c); }

}  // namespace sg
