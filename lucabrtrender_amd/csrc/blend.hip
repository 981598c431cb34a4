// Blend: one 256-thread workgroup per 16x16 tile, one thread per pixel, each
// wave64 owning one 8x8 quadrant of the tile. The tile's pairs are walked front
// to back in batches staged in LDS:
//   1. each thread stages one pair's record (one coalesced gather per pair, not
//      one per pixel) and tests it against the four quadrants with the exact
//      minimum of the Gaussian's power over each quadrant's pixel centres;
//   2. the batch is compacted into one index list per quadrant (ballot counts,
//      a four-wave prefix), so each wave walks only the pairs that can reach one
//      of its 64 pixels -- about 1.6 of a pair's 4 quadrants on the 3M-splat
//      scene, 2.5x fewer pixel-pair steps than walking the whole tile;
//   3. each wave walks its list with a branch-free body (a skipped pair
//      contributes alpha 0, which changes nothing) and leaves the walk once its
//      64 pixels are opaque; the workgroup stops between batches when all are.
// Tiles are taken in the order the host gives (heaviest first), so the longest
// sequential walks start at once instead of trailing the launch.
//
// Per pixel it is the reference's splatBlend walk (shaders/lrt/splat/
// splat_blend.slang:23-108, splat-only entry :110-114): skip past the cutoff,
// alpha = min(0.99, opacity * exp(-power / 2)), skip below 1/255, accumulate
// with the transmittance, stop below 1e-4, then linearise rgb / coverage. The
// Gaussian's exponent runs on the hardware exp2 and the quadratic form on
// FMAs; both move the result by a few ulps, which the image tolerances cover.
// Skipping a pair for a whole quadrant needs a test that never drops a pixel
// the per-pixel test would keep: the quadrant minimum is compared with the
// cutoff widened by a relative 1e-5, far above the float rounding of either.
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "splat_math.h"

namespace lrt::kern {
namespace {

using namespace lrt::dev;

constexpr uint32_t kBlock = kTile * kTile;
constexpr uint32_t kWaves = kBlock / 64;
constexpr float kExpScale = -0.72134752044448170368f;   // -log2(e) / 2

// The finished pixel: linearise rgb / coverage, the background behind, the
// depth (reference splat_blend.slang:89-107, splat-only).
__device__ __forceinline__ void finishPixel(const FrameParams& f, uint32_t px, uint32_t py, float r, float g, float b,
                                            float t, float depthSum, float depthAt, uint32_t last,
                                            float4* __restrict__ colour, float* __restrict__ depth,
                                            uint32_t* __restrict__ lastCount, float* __restrict__ finalT) {
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

template <uint32_t DEPTH_MODE, uint32_t BATCH>
__global__ __launch_bounds__(kBlock) void blendKernel(FrameParams f, uint32_t slots, const uint32_t* __restrict__ order,
                                                      const uint2* __restrict__ ranges,
                                                      const uint32_t* __restrict__ pairSlots,
                                                      const float4* __restrict__ proj,
                                                      float4* __restrict__ colour, float* __restrict__ depth,
                                                      uint32_t* __restrict__ lastCount,
                                                      float* __restrict__ finalT) {
    static_assert(BATCH == 64 || BATCH == 128 || BATCH == 256, "batch depth");
    __shared__ float4 sGeom[BATCH];     // centre.xy, cutoff, alpha
    __shared__ float4 sConic[BATCH];    // a, 2b, c, view z
    __shared__ float4 sColour[BATCH];   // rgb, 0
    __shared__ uint16_t sList[kWaves][BATCH];
    __shared__ uint32_t sCount[kWaves][kWaves];   // [quadrant][wave that found them]

    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u;
    const uint32_t wave = tid >> 6;
    const uint32_t tile = order != nullptr ? order[blockIdx.x] : blockIdx.x;
    if (tile >= f.tilesX * f.tilesY) return;   // an order entry that is not a tile: draw nothing
    const uint32_t tileX = tile % f.tilesX;
    const uint32_t tileY = tile / f.tilesX;
    // wave w: quadrant (w & 1, w >> 1); lane: 8x8 inside it
    const uint32_t px = tileX * kTile + (wave & 1u) * 8u + (lane & 7u);
    const uint32_t py = tileY * kTile + (wave >> 1) * 8u + (lane >> 3);
    const bool inside = px < f.width && py < f.height;
    const uint2 range = ranges[tile];
    const float pxf = static_cast<float>(px) + 0.5f;
    const float pyf = static_cast<float>(py) + 0.5f;
    // Each quadrant's pixel-centre box, for the staging threads' tests.
    float boxLo[kWaves][2], boxHi[kWaves][2];
#pragma unroll
    for (uint32_t q = 0; q < kWaves; ++q) {
        const uint32_t x0 = tileX * kTile + (q & 1u) * 8u, y0 = tileY * kTile + (q >> 1) * 8u;
        const uint32_t x1 = x0 + 8u < f.width ? x0 + 8u : f.width;
        const uint32_t y1 = y0 + 8u < f.height ? y0 + 8u : f.height;
        boxLo[q][0] = static_cast<float>(x0) + 0.5f;
        boxLo[q][1] = static_cast<float>(y0) + 0.5f;
        // An empty quadrant (past the frame's edge) gets an inverted box: never reached.
        boxHi[q][0] = x1 > x0 ? static_cast<float>(x1 - 1u) + 0.5f : -1e30f;
        boxHi[q][1] = y1 > y0 ? static_cast<float>(y1 - 1u) + 0.5f : -1e30f;
    }

    float r = 0.0f, g = 0.0f, b = 0.0f, t = 1.0f, depthSum = 0.0f, depthAt = -1.0f;
    uint32_t last = 0;
    bool active = inside;
    for (uint32_t batch = range.x; batch < range.y; batch += BATCH) {
        if (__syncthreads_count(active ? 1 : 0) == 0) break;
        // 1. stage and test
        uint32_t mask = 0;
        if (tid < BATCH && batch + tid < range.y) {
            const uint32_t slot = pairSlots[batch + tid];
            float4 screen = make_float4(0.0f, 0.0f, 0.0f, -1.0f), conic = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            float4 col = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (slot < slots) {
                const uint64_t s3 = static_cast<uint64_t>(slot) * 3;
                screen = proj[s3];
                conic = proj[s3 + 1];
                col = proj[s3 + 2];
            }
            sGeom[tid] = make_float4(screen.x, screen.y, screen.w, conic.w);
            sConic[tid] = make_float4(conic.x, conic.y + conic.y, conic.z, screen.z);
            sColour[tid] = col;
            const float reach = screen.w * 1.00001f + 1e-6f;
#pragma unroll
            for (uint32_t q = 0; q < kWaves; ++q) {
                if (boxHi[q][0] < 0.0f || boxHi[q][1] < 0.0f) continue;
                const float m = minPowerOverBox(conic.x, conic.y, conic.z, boxLo[q][0] - screen.x,
                                                boxLo[q][1] - screen.y, boxHi[q][0] - screen.x,
                                                boxHi[q][1] - screen.y);
                if (m <= reach) mask |= 1u << q;
            }
        }
        // 2. compact: list q gets, in pair order, the batch entries reaching quadrant q
        uint32_t before[kWaves];
        const uint64_t below = (1ull << lane) - 1ull;
#pragma unroll
        for (uint32_t q = 0; q < kWaves; ++q) {
            const uint64_t vote = __ballot((mask >> q) & 1u);
            before[q] = static_cast<uint32_t>(__popcll(vote & below));
            if (lane == 0) sCount[q][wave] = static_cast<uint32_t>(__popcll(vote));
        }
        __syncthreads();
#pragma unroll
        for (uint32_t q = 0; q < kWaves; ++q) {
            if ((mask >> q) & 1u) {
                uint32_t at = before[q];
                for (uint32_t w = 0; w < wave; ++w) at += sCount[q][w];
                sList[q][at] = static_cast<uint16_t>(tid);
            }
        }
        __syncthreads();
        // 3. walk this wave's list
        const uint32_t n = sCount[wave][0] + sCount[wave][1] + sCount[wave][2] + sCount[wave][3];
        if (__ballot(active) != 0) {
            for (uint32_t j = 0; j < n; ++j) {
                const uint32_t e = sList[wave][j];
                const float4 r0 = sGeom[e];
                const float4 r1 = sConic[e];
                const float dx = pxf - r0.x;
                const float dy = pyf - r0.y;
                const float power = __builtin_fmaf(dx, __builtin_fmaf(r1.x, dx, r1.y * dy), (r1.z * dy) * dy);
                const float alpha = fminS(0.99f, r0.w * __builtin_amdgcn_exp2f(kExpScale * power));
                const bool ok = active && power <= r0.z && alpha >= 1.0f / 255.0f;
                const float a = ok ? alpha : 0.0f;
                const float4 r2 = sColour[e];
                const float weight = t * a;
                r = __builtin_fmaf(weight, r2.x, r);
                g = __builtin_fmaf(weight, r2.y, g);
                b = __builtin_fmaf(weight, r2.z, b);
                depthSum = __builtin_fmaf(weight, r1.w, depthSum);
                t = t * (1.0f - a);
                last = ok ? batch + e - range.x + 1 : last;
                if constexpr (DEPTH_MODE == 1) {
                    depthAt = (ok && depthAt < 0.0f && 1.0f - t >= f.depthThreshold) ? r1.w : depthAt;
                }
                active = active && !(t < 1e-4f);
                if (__ballot(active) == 0) break;
            }
        }
    }
    if (!inside) return;
    finishPixel(f, px, py, r, g, b, t, depthSum, depthAt, last, colour, depth, lastCount, finalT);
}

// Heaviest tiles first: tiles bucketed by the bit length of their pair count,
// buckets in descending order, tiles in index order inside a bucket. One
// workgroup; bucket counts in LDS.
constexpr uint32_t kBuckets = 33;

__global__ __launch_bounds__(1024) void tileOrderKernel(const uint2* __restrict__ ranges, uint32_t tiles,
                                                        uint32_t* __restrict__ order) {
    __shared__ uint32_t sCount[kBuckets];
    const uint32_t tid = threadIdx.x;
    if (tid < kBuckets) sCount[tid] = 0;
    __syncthreads();
    for (uint32_t t = tid; t < tiles; t += blockDim.x) {
        const uint2 rg = ranges[t];
        const uint32_t bucket = 32u - static_cast<uint32_t>(__clz(rg.y - rg.x));   // 0 for an empty tile
        atomicAdd(&sCount[bucket], 1u);
    }
    __syncthreads();
    // Stable within a bucket: one wave walks the tiles in order. Lane k owns the
    // next free position of bucket k in a register and hands it out by shuffle:
    // positions never pass between lanes through memory, which the compiler
    // would be free to cache per lane.
    if (tid < 64) {
        uint32_t next = 0;
        if (tid < kBuckets) {
            for (uint32_t k = tid + 1; k < kBuckets; ++k) next += sCount[k];
        }
        const uint64_t below = (1ull << tid) - 1ull;
        for (uint32_t base = 0; base < tiles; base += 64) {
            const uint32_t t = base + tid;
            uint32_t bucket = kBuckets;
            if (t < tiles) {
                const uint2 rg = ranges[t];
                bucket = 32u - static_cast<uint32_t>(__clz(rg.y - rg.x));
            }
            // One round per bucket present among these 64 tiles.
            uint64_t todo = __ballot(t < tiles);
            while (todo != 0) {
                const int first = static_cast<int>(__builtin_ctzll(todo));
                const uint32_t k = static_cast<uint32_t>(__shfl(static_cast<int>(bucket), first, 64));
                const uint64_t vote = __ballot(t < tiles && bucket == k);
                const uint32_t start = static_cast<uint32_t>(__shfl(static_cast<int>(next), static_cast<int>(k), 64));
                if (t < tiles && bucket == k) order[start + static_cast<uint32_t>(__popcll(vote & below))] = t;
                if (tid == k) next += static_cast<uint32_t>(__popcll(vote));
                todo &= ~vote;
            }
        }
    }
}

template <uint32_t DEPTH_MODE>
void launchBlend(uint32_t batch, const FrameParams& f, uint32_t slots, const uint32_t* order, const uint2* ranges,
                 const uint32_t* pairSlots, const float4* proj, float4* colour, float* depth, uint32_t* lastCount,
                 float* finalT, hipStream_t stream) {
    const uint32_t tiles = f.tilesX * f.tilesY;
    if (batch <= 64) {
        blendKernel<DEPTH_MODE, 64><<<tiles, kBlock, 0, stream>>>(f, slots, order, ranges, pairSlots, proj, colour,
                                                                  depth, lastCount, finalT);
    } else if (batch <= 128) {
        blendKernel<DEPTH_MODE, 128><<<tiles, kBlock, 0, stream>>>(f, slots, order, ranges, pairSlots, proj, colour,
                                                                   depth, lastCount, finalT);
    } else {
        blendKernel<DEPTH_MODE, 256><<<tiles, kBlock, 0, stream>>>(f, slots, order, ranges, pairSlots, proj, colour,
                                                                   depth, lastCount, finalT);
    }
}

}   // namespace

void tileOrder(const uint2* ranges, uint32_t tiles, uint32_t* order, hipStream_t stream) {
    if (tiles == 0) return;
    tileOrderKernel<<<1, 1024, 0, stream>>>(ranges, tiles, order);
}

void blend(const FrameParams& f, uint32_t batch, uint32_t slots, const uint32_t* order, const uint2* ranges,
           const uint32_t* pairSlots, const float4* proj, float4* colour, float* depth, uint32_t* lastCount,
           float* finalT, hipStream_t stream) {
    if (f.tilesX * f.tilesY == 0) return;
    if (f.depthMode == 0) {
        launchBlend<0>(batch, f, slots, order, ranges, pairSlots, proj, colour, depth, lastCount, finalT, stream);
    } else {
        launchBlend<1>(batch, f, slots, order, ranges, pairSlots, proj, colour, depth, lastCount, finalT, stream);
    }
}

}   // namespace lrt::kern
