// Packing into the GpuSplats layout, host and device alike (the reference's
// packing.slang:26-50,104-113 and splat_decode.slang:39-87): fp16 halves
// rounded to nearest even after a +-65000 clamp, a rotation normalised and
// stored smallest-three in 10 bits, rounded to nearest. The clamp's max/min
// return their second operand on a NaN, as HLSL's clamp does, so a NaN packs
// as -65000.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace lrt::pack {

__host__ __device__ inline uint32_t bitsOf(float f) { return __builtin_bit_cast(uint32_t, f); }
__host__ __device__ inline float floatOf(uint32_t u) { return __builtin_bit_cast(float, u); }
__host__ __device__ inline float maxOf(float a, float b) { return a > b ? a : b; }
__host__ __device__ inline float minOf(float a, float b) { return a < b ? a : b; }

// A float's nearest half (ties to even); past the range, infinity; NaN kept quiet.
__host__ __device__ inline uint16_t floatToHalf(float f) {
    const uint32_t x = bitsOf(f);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const uint32_t a = x & 0x7FFFFFFFu;
    if (a >= 0x7F800000u) {
        return static_cast<uint16_t>(sign | 0x7C00u | (a > 0x7F800000u ? (0x200u | ((a >> 13) & 0x3FFu)) : 0u));
    }
    if (a >= 0x477FF000u) return static_cast<uint16_t>(sign | 0x7C00u);   // rounds past 65504
    if (a < 0x38800000u) {
        // Subnormal half: units of 2^-24, rounded to nearest even.
        return static_cast<uint16_t>(sign | static_cast<uint32_t>(__builtin_rintf(floatOf(a) * 16777216.0f)));
    }
    uint32_t h = (((a >> 23) - 112u) << 10) | ((a & 0x7FFFFFu) >> 13);
    const uint32_t rest = a & 0x1FFFu;
    if (rest > 0x1000u || (rest == 0x1000u && (h & 1u) != 0u)) h += 1u;
    return static_cast<uint16_t>(sign | h);
}

__host__ __device__ inline float halfSafe(float v) { return minOf(maxOf(v, -65000.0f), 65000.0f); }

__host__ __device__ inline uint32_t packHalves(float lo, float hi) {
    return static_cast<uint32_t>(floatToHalf(halfSafe(lo))) | (static_cast<uint32_t>(floatToHalf(halfSafe(hi))) << 16);
}

// (x, y, z, w), any length > 0: normalised, sign folded so the largest
// component is positive, its index in bits 0-1 and the other three in 10 bits
// each from bit 22 down.
__host__ __device__ inline uint32_t encodeQuaternion(const float in[4]) {
    const float len = __builtin_sqrtf(((in[0] * in[0] + in[1] * in[1]) + in[2] * in[2]) + in[3] * in[3]);
    float q[4] = {in[0] / len, in[1] / len, in[2] / len, in[3] / len};
    uint32_t largest = 0;
    float best = __builtin_fabsf(q[0]);
    for (uint32_t k = 1; k < 4; ++k) {
        if (__builtin_fabsf(q[k]) > best) {
            best = __builtin_fabsf(q[k]);
            largest = k;
        }
    }
    if (q[largest] < 0.0f) {
        for (float& v : q) v = -v;
    }
    uint32_t word = largest;
    uint32_t shift = 22;
    for (uint32_t k = 0; k < 4; ++k) {
        if (k == largest) continue;
        const float unit = minOf(maxOf(q[k] / 0.70710678f * 0.5f + 0.5f, 0.0f), 1.0f);
        word |= static_cast<uint32_t>(__builtin_rintf(unit * 1023.0f)) << shift;
        shift -= 10;
    }
    return word;
}

}   // namespace lrt::pack
