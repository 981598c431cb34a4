// Points as opaque discs in the splat pipeline: one thread per point. A disc
// is a record like a splat's -- centre, view z, an isotropic conic of its
// radius, cutoff 1, alpha 1 -- with colour.w = 1, which the composite blend
// reads as "opaque: draw it and stop". Its tile rectangle is that of the
// radius itself, its tile count the same exact count a splat's is, its depth
// key the same bits(z) >> 7, so discs and splats share one depth order.
// Follows the reference's pointsProject (shaders/lrt/splat/points_project.slang:
// 33-86) with pointRadius (shaders/lrt/points/point_frame.slang:39-44).
//
// Bandwidth: reads 16 B of position and 8 B of colour per point; writes 8 + 4 B
// per point and 48 B per drawn one.
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "splat_math.h"

namespace lrt::kern {
namespace {

using namespace lrt::dev;

constexpr uint32_t kBlock = kProjectBlock;

// The disc's radius in pixels, never below half a pixel: a point that projects
// smaller still covers the pixel it lands in.
__device__ __forceinline__ float pointRadius(const FrameParams& f, const PointParams& p, float viewZ) {
    const float r = p.sizeMode == 1 ? p.size * 0.5f
                                    : (f.orthographic ? p.size * 0.5f * f.focalX : p.size * 0.5f * f.focalX / viewZ);
    return fmaxS(r, 0.5f);
}

__global__ __launch_bounds__(kBlock) void projectPointsKernel(FrameParams f, PointParams style,
                                                              const float4* __restrict__ positions,
                                                              const uint2* __restrict__ colours,
                                                              float4* __restrict__ proj, uint2* __restrict__ rects,
                                                              uint32_t* __restrict__ touchedOut, DepthBlocks depth) {
    __shared__ uint32_t sDepth[2 * (kBlock / 64)];
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    const bool live = i < f.count;   // every lane stays for the wave's tile count
    const uint32_t slot = f.base + i;
    const uint2 emptyRect = make_uint2(1u | (1u << 16), 0u);
    const float4 p = live ? positions[i] : make_float4(0.0f, 0.0f, 0.0f, 1.0f);
    const float* m = f.m;
    const float tx = m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3];
    const float ty = m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7];
    const float tz = m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11];
    uint32_t rect[4] = {1u, 1u, 0u, 0u};
    float cx = 0.0f, cy = 0.0f, k = 0.0f;
    bool candidate = false;
    if (live && tz > f.nearZ && !(tz > f.farZ)) {
        const float radius = pointRadius(f, style, tz);
        const float invZ = f.orthographic ? 1.0f : 1.0f / tz;
        cx = f.focalX * tx * invZ + f.centreX;
        cy = f.focalY * ty * invZ + f.centreY;
        k = 1.0f / (radius * radius);
        tileRectOfExtent(f.width, f.height, cx, cy, radius, radius, rect, f.tileShift);
        candidate = rect[0] <= rect[2] && rect[1] <= rect[3];
    }
    const uint32_t touched = countTouchedTilesWave(f.width, f.height, cx, cy, k, 0.0f, k, 1.0f, rect, candidate,
                                                   f.tileShift);
    if (depth.entries != nullptr) {
        waveDepthRange(touched > 0, tz, sDepth);
        __syncthreads();
        storeDepthRange(depth, sDepth);
    }
    if (!live) return;
    if (touched == 0) {
        touchedOut[slot] = 0;
        rects[slot] = emptyRect;
        return;
    }
    float r = style.colour[0], g = style.colour[1], b = style.colour[2];
    if (!style.constantColour) {
        const uint2 c = colours[i];
        r = lowHalf(c.x);
        g = highHalf(c.x);
        b = lowHalf(c.y);
    }
    float4* out = proj + static_cast<uint64_t>(slot) * 3;
    out[0] = make_float4(cx, cy, tz, 1.0f);
    out[1] = make_float4(k, 0.0f, k, 1.0f);
    out[2] = make_float4(r, g, b, 1.0f);
    rects[slot] = make_uint2(rect[0] | (rect[1] << 16), rect[2] | (rect[3] << 16));
    touchedOut[slot] = touched;
}

}   // namespace

void projectPoints(const FrameParams& f, const PointParams& style, const float4* positions, const uint32_t* colours,
                   float4* proj, uint2* rects, uint32_t* touched, DepthBlocks depth, hipStream_t stream) {
    if (f.count == 0) return;
    projectPointsKernel<<<projectBlocks(f.count), kBlock, 0, stream>>>(
        f, style, positions, reinterpret_cast<const uint2*>(colours), proj, rects, touched, depth);
}

}   // namespace lrt::kern
