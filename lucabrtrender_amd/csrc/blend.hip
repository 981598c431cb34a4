// Blend: one workgroup per 16x16 tile, one thread per pixel. The tile's pairs
// are walked front to back in batches of 256: each thread stages one pair's
// record into LDS (one coalesced gather per pair instead of one per pixel),
// then every thread walks the batch from LDS. The workgroup stops when all of
// its pixels are opaque (transmittance below 1e-4), which it learns from a
// barrier count between batches.
//
// Per pixel it is the reference's splatBlend walk (shaders/lrt/splat/
// splat_blend.slang:23-108, splat-only entry :110-114): skip past the cutoff,
// alpha = min(0.99, opacity * exp(-power / 2)), skip below 1/255, accumulate
// with the transmittance, stop below 1e-4, then linearise rgb / coverage. The
// Gaussian's exponent runs on the hardware exp2 and the quadratic form on
// FMAs; both move the result by a few ulps, which the image tolerances cover.
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "splat_math.h"

namespace lrt::kern {
namespace {

using namespace lrt::dev;

constexpr uint32_t kBlock = kTile * kTile;

__global__ __launch_bounds__(kBlock) void blendKernel(FrameParams f, uint32_t slots, const uint2* __restrict__ ranges,
                                                      const uint32_t* __restrict__ pairSlots,
                                                      const float4* __restrict__ proj,
                                                      float4* __restrict__ colour, float* __restrict__ depth,
                                                      uint32_t* __restrict__ lastCount,
                                                      float* __restrict__ finalT) {
    __shared__ float4 sScreen[kBlock];
    __shared__ float4 sConic[kBlock];
    __shared__ float4 sColour[kBlock];
    const uint32_t tid = threadIdx.x;
    const uint32_t tile = blockIdx.x;
    const uint32_t tileX = tile % f.tilesX;
    const uint32_t tileY = tile / f.tilesX;
    const uint32_t px = tileX * kTile + (tid % kTile);
    const uint32_t py = tileY * kTile + (tid / kTile);
    const bool inside = px < f.width && py < f.height;
    const uint2 range = ranges[tile];
    const float pxf = static_cast<float>(px) + 0.5f;
    const float pyf = static_cast<float>(py) + 0.5f;
    const float kExpScale = -0.72134752044448170368f;   // -log2(e) / 2

    float r = 0.0f, g = 0.0f, b = 0.0f, t = 1.0f, depthSum = 0.0f, depthAt = -1.0f;
    uint32_t last = 0;
    bool done = !inside;
    for (uint32_t batch = range.x; batch < range.y; batch += kBlock) {
        if (__syncthreads_count(done ? 1 : 0) == static_cast<int>(kBlock)) break;
        const uint32_t k = batch + tid;
        if (k < range.y) {
            const uint32_t slot = pairSlots[k];
            if (slot < slots) {
                const uint64_t s = static_cast<uint64_t>(slot) * 3;
                sScreen[tid] = proj[s];
                sConic[tid] = proj[s + 1];
                sColour[tid] = proj[s + 2];
            } else {
                // Never reached with consistent inputs: an out-of-range slot draws nothing.
                sScreen[tid] = make_float4(0.0f, 0.0f, 0.0f, -1.0f);
                sConic[tid] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                sColour[tid] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            }
        }
        __syncthreads();
        const uint32_t n = range.y - batch < kBlock ? range.y - batch : kBlock;
        if (!done) {
            for (uint32_t j = 0; j < n; ++j) {
                const float4 sc = sScreen[j];
                const float4 co = sConic[j];
                const float dx = pxf - sc.x;
                const float dy = pyf - sc.y;
                const float inner = __builtin_fmaf(co.x, dx, (2.0f * co.y) * dy);
                const float power = __builtin_fmaf(dx, inner, (co.z * dy) * dy);
                if (power > sc.w) continue;
                const float alpha = fminS(0.99f, co.w * __builtin_amdgcn_exp2f(kExpScale * power));
                if (alpha < 1.0f / 255.0f) continue;
                const float4 cl = sColour[j];
                const float weight = t * alpha;
                r = __builtin_fmaf(weight, cl.x, r);
                g = __builtin_fmaf(weight, cl.y, g);
                b = __builtin_fmaf(weight, cl.z, b);
                depthSum = __builtin_fmaf(weight, sc.z, depthSum);
                t = t * (1.0f - alpha);
                last = batch + j - range.x + 1;
                if (depthAt < 0.0f && 1.0f - t >= f.depthThreshold) depthAt = sc.z;
                if (t < 1e-4f) {
                    done = true;
                    break;
                }
            }
        }
    }
    if (!inside) return;
    const float coverage = 1.0f - t;
    if (f.linearise && coverage > 1e-6f) {
        r = srgbToLinear(r / coverage) * coverage;
        g = srgbToLinear(g / coverage) * coverage;
        b = srgbToLinear(b / coverage) * coverage;
    }
    const uint32_t at = py * f.width + px;
    colour[at] = make_float4(r + t * f.bg[0], g + t * f.bg[1], b + t * f.bg[2], coverage + t * f.bg[3]);
    if (depth != nullptr) {
        depth[at] = f.depthMode == 0 ? (coverage > 1e-6f ? depthSum / coverage : 0.0f) : fmaxS(depthAt, 0.0f);
    }
    if (lastCount != nullptr) lastCount[at] = last;
    if (finalT != nullptr) finalT[at] = t;
}

}   // namespace

void blend(const FrameParams& f, uint32_t slots, const uint2* ranges, const uint32_t* pairSlots, const float4* proj, float4* colour,
           float* depth, uint32_t* lastCount, float* finalT, hipStream_t stream) {
    const uint32_t tiles = f.tilesX * f.tilesY;
    if (tiles == 0) return;
    blendKernel<<<tiles, kBlock, 0, stream>>>(f, slots, ranges, pairSlots, proj, colour, depth, lastCount, finalT);
}

}   // namespace lrt::kern
