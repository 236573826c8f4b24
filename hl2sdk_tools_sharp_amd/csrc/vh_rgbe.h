// The lightmap colour encode shared by the lump writers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#pragma clang fp contract(off)

namespace vh {

// StockLightColor.Encode: the largest channel (ties to the later one) sets
// the exponent, biased exponent - 134 (0 for a zero maximum); each channel is
// multiplied by 2^-exponent and truncated to a byte as .NET's saturating
// float -> int conversion does.
__device__ __forceinline__ int32_t to_int_sat(float x)
{
    if (x != x) return 0;
    if (x >= 2147483648.0f) return 2147483647;
    if (x <= -2147483648.0f) return (int32_t)0x80000000;
    return (int32_t)x;
}

__device__ __forceinline__ uint32_t encode_one(float r, float g, float b)
{
    float mx = r > g ? (r > b ? r : b) : (g > b ? g : b);
    int e = 0;
    if (mx != 0.0f) e = (int)((__float_as_uint(mx) & 0x7F800000u) >> 23) - (7 + 127);
    float scalar = __uint_as_float((uint32_t)(127 - e) << 23);
    uint32_t R = (uint8_t)to_int_sat(r * scalar);
    uint32_t G = (uint8_t)to_int_sat(g * scalar);
    uint32_t B = (uint8_t)to_int_sat(b * scalar);
    uint32_t E = (uint8_t)(int8_t)e;
    return R | (G << 8) | (B << 16) | (E << 24);
}

}  // namespace vh
