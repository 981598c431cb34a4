// Raw splat records into packed splats on the GPU: the reference's
// CloudLoader::decodeSlice (modules/scene/src/GpuClouds.cpp:151-185) with
// splat_validate.slang:16-29, a prefix sum and splat_decode.slang:17-87.
//
//   validate  one thread a record: finite position and log scales, |log s| < 60,
//             opacity >= 1/255 (the blend would discard every fragment below);
//   scan      the valid flags' exclusive sum: each kept record's place, in file
//             order;
//   decode    one thread a record: position and linear opacity, the rotation
//             normalised and stored smallest-three, f16 log scales and base
//             colour, the kept rest coefficients as f16 rgb per basis;
//   bounds    the kept positions' box, per-workgroup partials then one sum.
//
// Bandwidth: a record is read twice (4 * floatsPerRecord B each: 248 B for a
// degree-3 PLY) and written once packed (32 B + 4 * shWords, 124 B at degree 3);
// the validate pass reads only the eight fields it tests.
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "packing.h"
#include "splat_math.h"

namespace lrt::kern {
namespace {

using namespace lrt::dev;

constexpr uint32_t kBlock = 256;
constexpr float kShC0 = 0.28209479177387814f;

// splat_encoding.slang:28-96, the modes SplatEncoding names.
__device__ __forceinline__ float decodeOpacity(float v, uint32_t mode) {
    if (mode == 0) return 1.0f / (1.0f + expF(-v));
    if (mode == 2) return v / 255.0f;
    return v;
}
__device__ __forceinline__ float decodeLogScale(float v, uint32_t mode) {
    if (mode == 2) return v / 16.0f - 10.0f;   // SPZ byte
    return mode == 0 ? v : logF(fmaxS(v, 1e-30f));
}
__device__ __forceinline__ float decodeColour(float v, uint32_t mode) {
    if (mode == 0) return 0.5f + kShC0 * v;
    if (mode == 2) return v / 255.0f;
    if (mode == 3) return 0.5f + kShC0 * ((v / 255.0f - 0.5f) / 0.15f);   // SPZ byte: its DC scale 0.15
    return v;
}
__device__ __forceinline__ float byteUnit(float v, uint32_t mode) { return mode == 1 ? (v - 128.0f) / 128.0f : v; }

// A file stored right-up-back (SPZ) is read right-down-front, a half turn
// about x: y and z of positions and rotations change sign, and so does a rest
// coefficient whose basis is odd in y xor z (3DGS basis order).
__device__ __forceinline__ float restFlipSign(uint32_t basis) {
    constexpr uint32_t kNegative = 0b010110101001011u;   // bases 0 1 3 6 8 10 11 13 (bit = basis)
    return ((kNegative >> (basis < 14u ? basis : 14u)) & 1u) != 0u ? -1.0f : 1.0f;
}

__device__ __forceinline__ void decodePosition(const float* __restrict__ r, const DecodeParams& p, float v[3]) {
    v[0] = r[p.x] * p.positionScale;
    v[1] = r[p.y] * p.positionScale;
    v[2] = r[p.z] * p.positionScale;
    if (p.flipYZ != 0) {
        v[1] = -v[1];
        v[2] = -v[2];
    }
}

// The record's rotation as (x, y, z, w), not yet normalised
// (splat_encoding.slang decodeQuaternion).
__device__ __forceinline__ void decodeRotation(const float* __restrict__ r, const DecodeParams& p, float q[4]) {
    if (p.rotationMode == 2) {
        // SPZ v2: x y z bytes, w from unit length.
        q[0] = r[p.rotX] / 127.5f - 1.0f;
        q[1] = r[p.rotY] / 127.5f - 1.0f;
        q[2] = r[p.rotZ] / 127.5f - 1.0f;
        q[3] = __builtin_sqrtf(fmaxS(0.0f, 1.0f - ((q[0] * q[0] + q[1] * q[1]) + q[2] * q[2])));
    } else if (p.rotationMode == 3) {
        // SPZ v3: smallest three in one u32 held as two 16-bit halves: two bits
        // name the largest component, then the others from the last down, a
        // 9-bit magnitude of sqrt(1/2) and a sign bit each.
        uint32_t comp = static_cast<uint32_t>(r[p.rotX]) | (static_cast<uint32_t>(r[p.rotY]) << 16);
        const uint32_t largest = comp >> 30;
        float sum = 0.0f;
        for (int i = 3; i >= 0; --i) {
            if (static_cast<uint32_t>(i) == largest) {
                continue;
            }
            const float magnitude = 0.70710678f * static_cast<float>(comp & 511u) / 511.0f;
            q[i] = ((comp >> 9) & 1u) != 0u ? -magnitude : magnitude;
            sum = sum + q[i] * q[i];
            comp >>= 10;
        }
        q[largest] = __builtin_sqrtf(fmaxS(0.0f, 1.0f - sum));
    } else {
        q[0] = byteUnit(r[p.rotX], p.rotationMode);
        q[1] = byteUnit(r[p.rotY], p.rotationMode);
        q[2] = byteUnit(r[p.rotZ], p.rotationMode);
        q[3] = byteUnit(r[p.rotW], p.rotationMode);
    }
    if (p.flipYZ != 0) {
        q[1] = -q[1];
        q[2] = -q[2];
    }
}

__device__ __forceinline__ bool keepRecord(const float* __restrict__ r, const DecodeParams& p) {
    float pos[3];
    decodePosition(r, p, pos);
    const float px = pos[0], py = pos[1], pz = pos[2];
    const float a = decodeOpacity(r[p.opacity], p.opacityMode);
    const float l0 = decodeLogScale(r[p.scale0], p.scaleMode);
    const float l1 = decodeLogScale(r[p.scale1], p.scaleMode);
    const float l2 = decodeLogScale(r[p.scale2], p.scaleMode);
    const bool finite = __builtin_isfinite(px) && __builtin_isfinite(py) && __builtin_isfinite(pz) &&
                        __builtin_isfinite(l0) && __builtin_isfinite(l1) && __builtin_isfinite(l2) &&
                        __builtin_fabsf(l0) < 60.0f && __builtin_fabsf(l1) < 60.0f && __builtin_fabsf(l2) < 60.0f;
    return finite && a >= 1.0f / 255.0f;
}

__global__ __launch_bounds__(kBlock) void validateKernel(const float* __restrict__ raw, DecodeParams p,
                                                         uint32_t* __restrict__ valid) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= p.count) return;
    valid[i] = keepRecord(raw + static_cast<uint64_t>(i) * p.stride, p) ? 1u : 0u;
}

__global__ __launch_bounds__(kBlock) void decodeKernel(const float* __restrict__ raw, DecodeParams p,
                                                       const uint32_t* __restrict__ valid,
                                                       const uint32_t* __restrict__ dest, uint32_t base,
                                                       float4* __restrict__ positions, uint4* __restrict__ shape,
                                                       uint32_t* __restrict__ sh) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= p.count || valid[i] == 0) return;
    const float* r = raw + static_cast<uint64_t>(i) * p.stride;
    const uint32_t j = base + dest[i];
    const float a = decodeOpacity(r[p.opacity], p.opacityMode);
    float pos[3];
    decodePosition(r, p, pos);
    positions[j] = make_float4(pos[0], pos[1], pos[2], clampS(a, 0.0f, 1.0f));
    float q[4];
    decodeRotation(r, p, q);
    const float qLength = __builtin_sqrtf(((q[0] * q[0] + q[1] * q[1]) + q[2] * q[2]) + q[3] * q[3]);
    if (!(qLength > 1e-8f)) {
        q[0] = 0.0f;
        q[1] = 0.0f;
        q[2] = 0.0f;
        q[3] = 1.0f;
    }
    const float l0 = decodeLogScale(r[p.scale0], p.scaleMode);
    const float l1 = decodeLogScale(r[p.scale1], p.scaleMode);
    const float l2 = decodeLogScale(r[p.scale2], p.scaleMode);
    const float b0 = decodeColour(r[p.dc0], p.colourMode);
    const float b1 = decodeColour(r[p.dc1], p.colourMode);
    const float b2 = decodeColour(r[p.dc2], p.colourMode);
    shape[j] = make_uint4(pack::encodeQuaternion(q), pack::packHalves(l0, l1), pack::packHalves(l2, b0),
                          pack::packHalves(b1, b2));
    if (p.keep == 0) return;
    // Rest coefficients rgb per basis, two halves a word; the file's own
    // per-colour count says where green starts in a colour-outer record.
    uint32_t* row = sh + static_cast<uint64_t>(j) * p.shWords;
    uint32_t word = 0, filled = 0, at = 0;
    for (uint32_t basis = 0; basis < p.keep; ++basis) {
        for (uint32_t colour = 0; colour < 3; ++colour) {
            const uint32_t field = p.restColourOuter != 0 ? p.restBase + colour * p.filePerColour + basis
                                                          : p.restBase + basis * 3 + colour;
            float value = byteUnit(r[field], p.restMode);
            if (p.flipYZ != 0) value *= restFlipSign(basis);
            word |= static_cast<uint32_t>(pack::floatToHalf(pack::halfSafe(value))) << (filled * 16);
            if (++filled == 2) {
                row[at++] = word;
                word = 0;
                filled = 0;
            }
        }
    }
    if (filled == 1) row[at] = word;
}

// Per-workgroup min and max of x, y, z over [0, count): partial b holds six floats.
__global__ __launch_bounds__(kBlock) void boundsPartialKernel(const float4* __restrict__ positions, uint32_t count,
                                                              float* __restrict__ partial) {
    __shared__ float sBox[6][kBlock / 64];
    float lo[3] = {__builtin_inff(), __builtin_inff(), __builtin_inff()};
    float hi[3] = {-__builtin_inff(), -__builtin_inff(), -__builtin_inff()};
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < count; i += gridDim.x * kBlock) {
        const float4 v = positions[i];
        lo[0] = fminS(lo[0], v.x); lo[1] = fminS(lo[1], v.y); lo[2] = fminS(lo[2], v.z);
        hi[0] = fmaxS(hi[0], v.x); hi[1] = fmaxS(hi[1], v.y); hi[2] = fmaxS(hi[2], v.z);
    }
    for (uint32_t d = 32; d > 0; d >>= 1) {
        for (int k = 0; k < 3; ++k) {
            lo[k] = fminS(lo[k], __shfl_xor(lo[k], static_cast<int>(d), 64));
            hi[k] = fmaxS(hi[k], __shfl_xor(hi[k], static_cast<int>(d), 64));
        }
    }
    const uint32_t wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63u) == 0) {
        for (int k = 0; k < 3; ++k) {
            sBox[k][wave] = lo[k];
            sBox[3 + k][wave] = hi[k];
        }
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        float v = sBox[threadIdx.x][0];
        for (uint32_t w = 1; w < kBlock / 64; ++w) {
            v = threadIdx.x < 3 ? fminS(v, sBox[threadIdx.x][w]) : fmaxS(v, sBox[threadIdx.x][w]);
        }
        partial[blockIdx.x * 6 + threadIdx.x] = v;
    }
}

__global__ void boundsFinishKernel(const float* __restrict__ partial, uint32_t blocks, float* __restrict__ box) {
    const uint32_t k = threadIdx.x;
    if (k >= 6) return;
    float v = partial[k];
    for (uint32_t b = 1; b < blocks; ++b) v = k < 3 ? fminS(v, partial[b * 6 + k]) : fmaxS(v, partial[b * 6 + k]);
    box[k] = v;
}

}   // namespace

uint32_t boundsPartials(uint32_t count) {
    const uint32_t blocks = (count + kBlock - 1) / kBlock;
    return blocks < 1024 ? (blocks > 0 ? blocks : 1u) : 1024u;
}

void decodeSplats(const float* raw, const DecodeParams& p, uint32_t* valid, uint32_t* dest, uint64_t* scanStatus,
                  uint32_t* kept, uint32_t* nonzero, uint32_t* scanError, uint32_t base, float4* positions,
                  uint4* shape, uint32_t* sh, hipStream_t stream) {
    if (p.count == 0) {
        fill32(kept, 0u, 1, stream);
        return;
    }
    const uint32_t blocks = (p.count + kBlock - 1) / kBlock;
    validateKernel<<<blocks, kBlock, 0, stream>>>(raw, p, valid);
    scanExclusive(valid, dest, p.count, scanStatus, kept, nonzero, scanError, stream);
    decodeKernel<<<blocks, kBlock, 0, stream>>>(raw, p, valid, dest, base, positions, shape, sh);
}

void splatBounds(const float4* positions, uint32_t count, float* partial, float* box, hipStream_t stream) {
    const uint32_t blocks = boundsPartials(count);
    boundsPartialKernel<<<blocks, kBlock, 0, stream>>>(positions, count, partial);
    boundsFinishKernel<<<1, 64, 0, stream>>>(partial, blocks, box);
}

}   // namespace lrt::kern
