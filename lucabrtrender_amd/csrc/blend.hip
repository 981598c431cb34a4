// Blend: one wave64 workgroup per 8x8 quadrant of a 16x16 tile, one lane per
// pixel. The wave walks its tile's sorted pairs front to back in rounds of 128:
//   1. each lane takes two pairs of the round (their records gathered one round
//      ahead, so the loads are in flight while the previous round is worked)
//      and tests each against the quadrant's pixel-centre box with a
//      conservative bound on the Gaussian's power (quadrantReach);
//   2. the pairs that pass are compacted (ballot + popcount) into the wave's
//      own LDS list, records and all, in pair order, until the list is full or
//      the pairs run out;
//   3. the wave walks the list kGroup records at a time with a branch-free body
//      (a skipped pair contributes alpha 0, which changes nothing) and leaves
//      once its 64 pixels are opaque.
// A staged record carries its Gaussian's form expanded about the quadrant's
// centre and scaled to the exponent's base 2, the cutoff folded in: a pixel at
// (u, v) from the centre evaluates q = A u^2 + B u v + C v^2 + D u + E v + F
// with five FMAs, passes the cutoff test when q >= 0, and has alpha =
// min(0.99, o' 2^q), o' = opacity * 2^(cutoff scaled). q differs from the
// direct form's scaled power by float rounding only (below 2e-4 on footprints
// from a quarter pixel to 300 pixels wide, centres up to 3.5 sigma outside the
// quadrant).
// No wave ever waits for another: quadrants finish as soon as their own pixels
// do, and a quadrant with few pairs frees its slot for the next. The four
// quadrants of one tile are blocks b, b + 8, b + 16, b + 24, which the
// dispatcher deals to one XCD, so their gathers of the tile's records share
// that XCD's L2.
//
// Per pixel it is the reference's splatBlend walk (shaders/lrt/splat/
// splat_blend.slang:23-108, splat-only entry :110-114): skip past the cutoff,
// alpha = min(0.99, opacity * exp(-power / 2)), skip below 1/255, accumulate
// with the transmittance, stop below 1e-4, then linearise rgb / coverage. The
// Gaussian's exponent runs on the hardware exp2 and the quadratic form on
// FMAs; both move the result by a few ulps, which the image tolerances cover.
//
// The composite walk (splatBlendComposite, :116-120) is the same kernel
// instantiated with COMPOSITE: a disc record (colour.w 1) that reaches a pixel
// ends its walk and lies behind what it blended; a pixel with an under-layer
// depth above 0 ends its walk at the first pair of its tile at or behind that
// depth. That first pair may be one the quadrant's list dropped, so each staged
// record carries the largest view z of every pair up to it (a wave max-scan at
// staging), and the walk tests that instead of the record's own z.
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "splat_math.h"

namespace lrt::kern {
namespace {

using namespace lrt::dev;

constexpr uint32_t kLanes = 64;          // one wave: an 8x8 quadrant
constexpr uint32_t kPerLane = 2;         // pairs a lane stages per round
constexpr uint32_t kRound = kLanes * kPerLane;
constexpr uint32_t kGroup = 4;           // list records walked per step
constexpr float kExpScale = -0.72134752044448170368f;   // -log2(e) / 2

// The finished pixel: linearise rgb / coverage, the background behind, the
// depth (reference splat_blend.slang:89-107, splat-only).
__device__ __forceinline__ void finishPixel(const FrameParams& f, uint32_t px, uint32_t py, float r, float g, float b,
                                            float t, float depthSum, float depthAt, uint32_t last,
                                            float4* __restrict__ colour, float* __restrict__ depth,
                                            uint32_t* __restrict__ lastCount, float* __restrict__ finalT,
                                            float4* __restrict__ accum) {
    const float coverage = 1.0f - t;
    if (accum != nullptr) accum[py * f.width + px] = make_float4(r, g, b, depthSum);
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

// The finished pixel of the composite walk (reference splat_blend.slang:83-107):
// without a disc, the under-layer lies behind; whatever lies behind is
// composited under the linearised splats and weighs in the depth with the
// transmittance left.
__device__ __forceinline__ void finishComposite(const FrameParams& f, uint32_t at, float r, float g, float b, float t,
                                                float depthSum, float depthAt, bool opaqueHit, float4 opaque,
                                                float opaqueZ, bool underHit, const float4* __restrict__ underColour,
                                                float underZ, float4* __restrict__ colour, float* __restrict__ depth) {
    if (!opaqueHit && underHit) {
        opaqueHit = true;
        opaque = underColour[at];
        opaqueZ = underZ;
    }
    const float coverage = 1.0f - t;
    if (f.linearise && coverage > 1e-6f) {
        r = srgbToLinear(r / coverage) * coverage;
        g = srgbToLinear(g / coverage) * coverage;
        b = srgbToLinear(b / coverage) * coverage;
    }
    float4 behind = make_float4(f.bg[0], f.bg[1], f.bg[2], f.bg[3]);
    if (opaqueHit) {
        behind = opaque;
        depthSum = depthSum + t * opaqueZ;
        if (depthAt < 0.0f) depthAt = opaqueZ;
    }
    const float total = opaqueHit ? 1.0f : coverage;
    colour[at] = make_float4(r + t * behind.x, g + t * behind.y, b + t * behind.z, coverage + t * behind.w);
    if (depth != nullptr) {
        depth[at] = f.depthMode == 0 ? (total > 1e-6f ? depthSum / total : 0.0f) : fmaxS(depthAt, 0.0f);
    }
}

// One pair's projected record, or an empty one (cutoff -1) past the range.
struct Staged {
    float4 screen;   // centre.xy, view z, cutoff
    float4 conic;    // a, b, c, alpha
    float4 col;      // rgb, 0
};

__device__ __forceinline__ Staged gather(const float4* __restrict__ proj, uint32_t slot, uint32_t slots) {
    Staged s{make_float4(0.0f, 0.0f, 0.0f, -1.0f), make_float4(0.0f, 0.0f, 0.0f, 0.0f),
             make_float4(0.0f, 0.0f, 0.0f, 0.0f)};
    if (slot < slots) {
        const uint64_t s3 = static_cast<uint64_t>(slot) * 3;
        s.screen = proj[s3];
        s.conic = proj[s3 + 1];
        s.col = proj[s3 + 2];
    }
    return s;
}

// Quadrant `quadrant` of tile `tile` of the frame f, whose pairs are
// [range.x, range.y) of pairSlots; colour and depth are the frame's images,
// lastCount/finalT/accum its per-pixel state (each may be null).
template <uint32_t DEPTH_MODE, uint32_t LIST, bool COMPOSITE>
__device__ __forceinline__ void blendQuadrant(const FrameParams& f, uint32_t tile, uint32_t quadrant, uint2 range,
                                              uint32_t slots, const uint32_t* __restrict__ pairSlots,
                                              const float4* __restrict__ proj,
                                              const float4* __restrict__ underColour,
                                              const float* __restrict__ underDepth, float4* __restrict__ colour,
                                              float* __restrict__ depth, uint32_t* __restrict__ lastCount,
                                              float* __restrict__ finalT, float4* __restrict__ accum) {
    static_assert(LIST % kRound == 0 && LIST >= kRound, "list length");
    // The list: records that reach this quadrant, in pair order, padded to a
    // multiple of kGroup with empty records (cutoff -1: never pass the test).
    __shared__ float4 sGeom[LIST + kGroup];     // A, B, C, D of the expanded form
    __shared__ float4 sConic[LIST + kGroup];    // E, F less the cutoff, o', view z
    // rgb, then bits of the pair's place in the tile's range + 1; composite:
    // the largest view z of the tile's pairs up to this one, kept or not
    __shared__ float4 sColour[LIST + kGroup];

    const uint32_t lane = threadIdx.x;
    const uint32_t x0 = (tile % f.tilesX) * kTile + (quadrant & 1u) * 8u;
    const uint32_t y0 = (tile / f.tilesX) * kTile + (quadrant >> 1) * 8u;
    if (x0 >= f.width || y0 >= f.height) return;   // a quadrant past the frame's edge has no pixels
    const uint32_t x1 = x0 + 8u < f.width ? x0 + 8u : f.width;
    const uint32_t y1 = y0 + 8u < f.height ? y0 + 8u : f.height;
    const uint32_t px = x0 + (lane & 7u);
    const uint32_t py = y0 + (lane >> 3);
    const bool inside = px < x1 && py < y1;
    // The pixel centre about the quadrant's centre, and that centre.
    const float pu = static_cast<float>(lane & 7u) - 3.5f;
    const float pv = static_cast<float>(lane >> 3) - 3.5f;
    const float xm = static_cast<float>(x0) + 4.0f, ym = static_cast<float>(y0) + 4.0f;
    const float boxLoX = static_cast<float>(x0) + 0.5f, boxLoY = static_cast<float>(y0) + 0.5f;
    const float boxHiX = static_cast<float>(x1 - 1u) + 0.5f, boxHiY = static_cast<float>(y1 - 1u) + 0.5f;
    const uint64_t below = (1ull << lane) - 1ull;

    // Round k's records sit in `staged` while round k + 1's are gathered.
    auto slotAt = [&](uint32_t at) -> uint32_t { return at < range.y ? pairSlots[at] : 0xFFFFFFFFu; };
    Staged staged[kPerLane];
#pragma unroll
    for (uint32_t u = 0; u < kPerLane; ++u) staged[u] = gather(proj, slotAt(range.x + u * kLanes + lane), slots);

    float r = 0.0f, g = 0.0f, b = 0.0f, t = 1.0f, depthSum = 0.0f, depthAt = -1.0f;
    uint32_t last = 0;
    bool active = inside;
    // Composite: what ends the walk from behind. A pixel with an under depth
    // above 0 stops at the first pair at or behind it (whether or not that pair
    // reaches the pixel: the reference tests the depth first), a disc that
    // reaches the pixel stops it with its colour.
    const uint32_t at0 = py * f.width + px;
    float underZ = 0.0f;
    if constexpr (COMPOSITE) {
        if (underDepth != nullptr && inside) underZ = underDepth[at0];
    }
    const bool underHit = COMPOSITE && underZ > 0.0f;
    bool opaqueHit = false;
    float oR = 0.0f, oG = 0.0f, oB = 0.0f, oZ = 0.0f;
    float zCarry = -3.4e38f;   // the largest view z of the pairs staged so far
    uint32_t next = range.x;   // the first pair of the round in `staged`
    while (next < range.y) {
        // 1-2. stage rounds into the list while a whole round still fits
        uint32_t fill = 0;
        while (next < range.y && fill + kRound <= LIST) {
            Staged ahead[kPerLane];
#pragma unroll
            for (uint32_t u = 0; u < kPerLane; ++u) {
                ahead[u] = gather(proj, slotAt(next + kRound + u * kLanes + lane), slots);
            }
#pragma unroll
            for (uint32_t u = 0; u < kPerLane; ++u) {
                const Staged& s = staged[u];
                const uint32_t at = next + u * kLanes + lane;
                const bool keep = at < range.y &&
                                  quadrantReach(s.conic.x, s.conic.y, s.conic.z, s.screen.w, boxLoX - s.screen.x,
                                                boxLoY - s.screen.y, boxHiX - s.screen.x, boxHiY - s.screen.y);
                const uint64_t vote = __ballot(keep);
                float zReach = 0.0f;
                if constexpr (COMPOSITE) {
                    // inclusive max-scan of the round's view z in pair order
                    float zr = at < range.y ? s.screen.z : -3.4e38f;
#pragma unroll
                    for (uint32_t d = 1; d < kLanes; d <<= 1) {
                        const float o = __shfl_up(zr, d, 64);
                        zr = lane >= d ? fmaxS(zr, o) : zr;
                    }
                    zReach = fmaxS(zr, zCarry);
                    zCarry = shflF(zReach, kLanes - 1);
                }
                if (keep) {
                    const uint32_t to = fill + static_cast<uint32_t>(__popcll(vote & below));
                    const float ux = s.screen.x - xm, vy = s.screen.y - ym;
                    const float fa = kExpScale * s.conic.x, fb = kExpScale * (s.conic.y + s.conic.y);
                    const float fc = kExpScale * s.conic.z;
                    const float scaledCut = kExpScale * s.screen.w;
                    const float fd = -(2.0f * fa * ux + fb * vy), fe = -(2.0f * fc * vy + fb * ux);
                    const float ff = (ux * (fa * ux + fb * vy) + fc * vy * vy) - scaledCut;
                    sGeom[to] = make_float4(fa, fb, fc, fd);
                    // a disc (colour.w 1, composite only) is marked by o' < 0: its alpha never passes
                    const bool disc = COMPOSITE && s.col.w > 0.5f;
                    const float scaledOpacity = disc ? -1.0f : s.conic.w * __builtin_amdgcn_exp2f(scaledCut);
                    sConic[to] = make_float4(fe, ff, scaledOpacity, s.screen.z);
                    sColour[to] = make_float4(s.col.x, s.col.y, s.col.z,
                                              COMPOSITE ? zReach : __uint_as_float(at - range.x + 1u));
                }
                fill += static_cast<uint32_t>(__popcll(vote));
            }
#pragma unroll
            for (uint32_t u = 0; u < kPerLane; ++u) staged[u] = ahead[u];
            next += kRound;
        }
        const uint32_t padded = (fill + kGroup - 1u) & ~(kGroup - 1u);
        if (lane < padded - fill) {
            sGeom[fill + lane] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            sConic[fill + lane] = make_float4(0.0f, -1e30f, 0.0f, 0.0f);   // q far below 0: never passes
            sColour[fill + lane] = make_float4(0.0f, 0.0f, 0.0f, COMPOSITE ? -3.4e38f : 0.0f);
        }
        __syncthreads();
        // 3. walk the list: all kGroup records' reads issued together, then the
        // pixels blend them in order
        for (uint32_t j = 0; j < padded; j += kGroup) {
            float4 g0[kGroup], g1[kGroup], g2[kGroup];
#pragma unroll
            for (uint32_t k = 0; k < kGroup; ++k) {
                g0[k] = sGeom[j + k];
                g1[k] = sConic[j + k];
                g2[k] = sColour[j + k];
                // the place is needed only when the pair contributes; without
                // this the compiler loads it behind a branch in the walk
                asm volatile("" ::"v"(g2[k].w));
            }
#pragma unroll
            for (uint32_t k = 0; k < kGroup; ++k) {
                if constexpr (COMPOSITE) active = active & !(underHit & (g2[k].w >= underZ));
                const float q = __builtin_fmaf(pu, __builtin_fmaf(g0[k].x, pu, __builtin_fmaf(g0[k].y, pv, g0[k].w)),
                                               __builtin_fmaf(pv, __builtin_fmaf(g0[k].z, pv, g1[k].x), g1[k].y));
                const float alpha = __builtin_fminf(0.99f, g1[k].z * __builtin_amdgcn_exp2f(q));
                // bitwise, not short-circuit: no branch inside the walk
                const bool ok = active & (q >= 0.0f) & (alpha >= 1.0f / 255.0f);
                const float a = ok ? alpha : 0.0f;
                const float weight = t * a;
                r = __builtin_fmaf(weight, g2[k].x, r);
                g = __builtin_fmaf(weight, g2[k].y, g);
                b = __builtin_fmaf(weight, g2[k].z, b);
                depthSum = __builtin_fmaf(weight, g1[k].w, depthSum);
                t = t - weight;   // T (1 - alpha)
                if constexpr (!COMPOSITE) last = ok ? __float_as_uint(g2[k].w) : last;
                if constexpr (DEPTH_MODE == 1) {
                    depthAt = (ok && depthAt < 0.0f && 1.0f - t >= f.depthThreshold) ? g1[k].w : depthAt;
                }
                if constexpr (COMPOSITE) {
                    const bool hit = active & (q >= 0.0f) & (g1[k].z < 0.0f);
                    oR = hit ? g2[k].x : oR;
                    oG = hit ? g2[k].y : oG;
                    oB = hit ? g2[k].z : oB;
                    oZ = hit ? g1[k].w : oZ;
                    opaqueHit = opaqueHit | hit;
                    active = active & !hit;
                }
                active = active && !(t < 1e-4f);
            }
            if (__ballot(active) == 0) break;
        }
        if (__ballot(active) == 0) break;
        __syncthreads();   // the list is refilled next
    }
    if (!inside) return;
    if constexpr (COMPOSITE) {
        finishComposite(f, at0, r, g, b, t, depthSum, depthAt, opaqueHit, make_float4(oR, oG, oB, 1.0f), oZ,
                        underHit, underColour, underZ, colour, depth);
    } else {
        finishPixel(f, px, py, r, g, b, t, depthSum, depthAt, last, colour, depth, lastCount, finalT, accum);
    }
}

template <uint32_t DEPTH_MODE, uint32_t LIST, bool COMPOSITE>
__global__ __launch_bounds__(kLanes) void blendKernel(FrameParams f, uint32_t slots, const uint32_t* __restrict__ order,
                                                      const uint2* __restrict__ ranges,
                                                      const uint32_t* __restrict__ pairSlots,
                                                      const float4* __restrict__ proj,
                                                      const float4* __restrict__ underColour,
                                                      const float* __restrict__ underDepth,
                                                      float4* __restrict__ colour, float* __restrict__ depth,
                                                      uint32_t* __restrict__ lastCount,
                                                      float* __restrict__ finalT, float4* __restrict__ accum) {
    const uint32_t tiles = f.tilesX * f.tilesY;
    const uint32_t quadrant = (blockIdx.x >> 3) & 3u;
    const uint32_t place = (blockIdx.x >> 5) * 8u + (blockIdx.x & 7u);
    if (place >= tiles) return;
    const uint32_t tile = order != nullptr ? order[place] : place;
    if (tile >= tiles) return;   // an order entry that is not a tile: draw nothing
    blendQuadrant<DEPTH_MODE, LIST, COMPOSITE>(f, tile, quadrant, ranges[tile], slots, pairSlots, proj, underColour,
                                               underDepth, colour, depth, lastCount, finalT, accum);
}

// The view holding batch tile g: the last view whose tiles start at or before it.
__device__ __forceinline__ uint32_t viewOfTile(const ViewInfo* __restrict__ views, uint32_t viewCount, uint32_t g) {
    uint32_t lo = 0, hi = viewCount;   // views[lo].tileBase <= g < views[hi].tileBase
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (views[mid].tileBase <= g) lo = mid; else hi = mid;
    }
    return lo;
}

// Every quadrant of every view of a batch. As a grid (PERSISTENT false) block b
// is task b, the four quadrants of a tile going to blocks b, b + 8, b + 16,
// b + 24 of one XCD; persistent, each block takes tasks from a shared counter
// until none are left (task t: tile t / 4, quadrant t % 4), so a block that
// drew a light quadrant takes the next instead of leaving its slot idle.
template <uint32_t DEPTH_MODE, uint32_t LIST, bool PERSISTENT>
__global__ __launch_bounds__(kLanes) void blendViewsKernel(const FrameParams* __restrict__ params,
                                                           const ViewInfo* __restrict__ views, uint32_t viewCount,
                                                           uint32_t tiles, uint32_t tasks, uint32_t slots,
                                                           const uint2* __restrict__ ranges,
                                                           const uint32_t* __restrict__ pairSlots,
                                                           const float4* __restrict__ proj,
                                                           uint32_t* __restrict__ lastCount,
                                                           float* __restrict__ finalT, uint32_t* __restrict__ queue) {
    for (uint32_t round = 0;; ++round) {
        uint32_t task, g, quadrant;
        if constexpr (PERSISTENT) {
            uint32_t taken = 0;
            if (threadIdx.x == 0) taken = atomicAdd(queue, 1u);
            task = static_cast<uint32_t>(__shfl(static_cast<int>(taken), 0, 64));
            if (task >= tasks) return;
            g = task >> 2;
            quadrant = task & 3u;
            if (round > 0) __syncthreads();   // the last task's LDS list is done with
        } else {
            task = blockIdx.x;
            g = (task >> 5) * 8u + (task & 7u);
            quadrant = (task >> 3) & 3u;
        }
        if (g < tiles) {
            const ViewInfo& v = views[viewOfTile(views, viewCount, g)];
            blendQuadrant<DEPTH_MODE, LIST, false>(params[v.params], g - v.tileBase, quadrant, ranges[g], slots,
                                                   pairSlots, proj, nullptr, nullptr,
                                                   reinterpret_cast<float4*>(v.colour), v.depth,
                                            lastCount + v.pixelBase, finalT + v.pixelBase, nullptr);
        }
        if constexpr (!PERSISTENT) return;
    }
}

// Heaviest tiles first: tiles bucketed by the bit length of their pair count,
// buckets in descending order, tiles in index order inside a bucket. One
// workgroup of 16 waves, wave w owning the w-th sixteenth of the tiles: each
// wave counts its segment's buckets, the counts' prefix over (bucket
// descending, wave ascending) gives each wave its first place in every bucket,
// and each wave walks its segment again handing out places.
constexpr uint32_t kBuckets = 33;
constexpr uint32_t kOrderWaves = 16;
constexpr uint32_t kOrderCached = 32768;   // tiles whose bucket the first walk leaves in LDS for the second

__device__ __forceinline__ uint32_t bucketOf(uint2 rg) {
    return 32u - static_cast<uint32_t>(__clz(rg.y - rg.x));   // 0 for an empty tile
}

__global__ __launch_bounds__(kOrderWaves * 64) void tileOrderKernel(const uint2* __restrict__ ranges, uint32_t tiles,
                                                                    uint32_t* __restrict__ order) {
    __shared__ uint32_t sCount[kOrderWaves][kBuckets];
    __shared__ uint32_t sStart[kOrderWaves][kBuckets];
    __shared__ uint8_t sBucket[kOrderCached];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    for (uint32_t k = tid; k < kOrderWaves * kBuckets; k += blockDim.x) (&sCount[0][0])[k] = 0;
    __syncthreads();
    const uint32_t segment = (tiles + kOrderWaves * 64 - 1) / (kOrderWaves * 64) * 64;   // whole chunks of 64
    const uint32_t begin = wave * segment;
    const uint32_t end = begin + segment < tiles ? begin + segment : tiles;
    for (uint32_t t = begin + lane; t < end; t += 64) {
        const uint32_t bucket = bucketOf(ranges[t]);
        if (t < kOrderCached) sBucket[t] = static_cast<uint8_t>(bucket);
        atomicAdd(&sCount[wave][bucket], 1u);
    }
    __syncthreads();
    for (uint32_t e = tid; e < kOrderWaves * kBuckets; e += blockDim.x) {
        const uint32_t w = e / kBuckets, k = e % kBuckets;
        uint32_t start = 0;
        for (uint32_t kk = k + 1; kk < kBuckets; ++kk)
            for (uint32_t ww = 0; ww < kOrderWaves; ++ww) start += sCount[ww][kk];
        for (uint32_t ww = 0; ww < w; ++ww) start += sCount[ww][k];
        sStart[w][k] = start;
    }
    __syncthreads();
    // Lane k owns the next free place of bucket k in a register and hands it
    // out by shuffle: places never pass between lanes through memory, which the
    // compiler would be free to cache per lane.
    uint32_t next = lane < kBuckets ? sStart[wave][lane] : 0u;
    const uint64_t below = (1ull << lane) - 1ull;
    for (uint32_t base = begin; base < end; base += 64) {
        const uint32_t t = base + lane;
        const bool mine = t < end;
        const uint32_t bucket = !mine ? kBuckets : t < kOrderCached ? sBucket[t] : bucketOf(ranges[t]);
        // One round per bucket present among these 64 tiles.
        uint64_t todo = __ballot(mine);
        while (todo != 0) {
            const int first = static_cast<int>(__builtin_ctzll(todo));
            const uint32_t k = static_cast<uint32_t>(__shfl(static_cast<int>(bucket), first, 64));
            const uint64_t vote = __ballot(mine && bucket == k);
            const uint32_t start = static_cast<uint32_t>(__shfl(static_cast<int>(next), static_cast<int>(k), 64));
            if (mine && bucket == k) order[start + static_cast<uint32_t>(__popcll(vote & below))] = t;
            if (lane == k) next += static_cast<uint32_t>(__popcll(vote));
            todo &= ~vote;
        }
    }
}

template <uint32_t DEPTH_MODE, bool COMPOSITE>
void launchBlend(uint32_t list, const FrameParams& f, uint32_t slots, const uint32_t* order, const uint2* ranges,
                 const uint32_t* pairSlots, const float4* proj, const float4* underColour, const float* underDepth,
                 float4* colour, float* depth, uint32_t* lastCount, float* finalT, float4* accum,
                 hipStream_t stream) {
    // Four blocks a tile, the grid rounded up to whole groups of eight tiles.
    const uint32_t tiles = f.tilesX * f.tilesY;
    const uint32_t blocks = (tiles + 7u) / 8u * 32u;
    if (list <= 128) {
        blendKernel<DEPTH_MODE, 128, COMPOSITE><<<blocks, kLanes, 0, stream>>>(
            f, slots, order, ranges, pairSlots, proj, underColour, underDepth, colour, depth, lastCount, finalT,
            accum);
    } else {
        blendKernel<DEPTH_MODE, 256, COMPOSITE><<<blocks, kLanes, 0, stream>>>(
            f, slots, order, ranges, pairSlots, proj, underColour, underDepth, colour, depth, lastCount, finalT,
            accum);
    }
}

template <uint32_t DEPTH_MODE, uint32_t LIST>
void launchBlendViews(bool persistent, uint32_t gridCap, const FrameParams* params, const ViewInfo* views,
                      uint32_t viewCount, uint32_t tiles, uint32_t slots, const uint2* ranges, const uint32_t* pairSlots,
                      const float4* proj, uint32_t* lastCount, float* finalT, uint32_t* queue, hipStream_t stream) {
    if (persistent) {
        const uint32_t tasks = tiles * 4u;
        const uint32_t grid = tasks < gridCap ? tasks : gridCap;
        fill32(queue, 0u, 1, stream);
        blendViewsKernel<DEPTH_MODE, LIST, true><<<grid, kLanes, 0, stream>>>(
            params, views, viewCount, tiles, tasks, slots, ranges, pairSlots, proj, lastCount, finalT, queue);
    } else {
        const uint32_t tasks = (tiles + 7u) / 8u * 32u;
        blendViewsKernel<DEPTH_MODE, LIST, false><<<tasks, kLanes, 0, stream>>>(
            params, views, viewCount, tiles, tasks, slots, ranges, pairSlots, proj, lastCount, finalT, queue);
    }
}

}   // namespace

void blendViews(uint32_t depthMode, uint32_t list, bool persistent, uint32_t gridCap, const FrameParams* params,
                const ViewInfo* views, uint32_t viewCount, uint32_t tiles, uint32_t slots, const uint2* ranges,
                const uint32_t* pairSlots, const float4* proj, uint32_t* lastCount, float* finalT, uint32_t* queue,
                hipStream_t stream) {
    if (tiles == 0 || viewCount == 0) return;
    if (depthMode == 0) {
        if (list <= 128) {
            launchBlendViews<0, 128>(persistent, gridCap, params, views, viewCount, tiles, slots, ranges, pairSlots,
                                     proj, lastCount, finalT, queue, stream);
        } else {
            launchBlendViews<0, 256>(persistent, gridCap, params, views, viewCount, tiles, slots, ranges, pairSlots,
                                     proj, lastCount, finalT, queue, stream);
        }
    } else {
        if (list <= 128) {
            launchBlendViews<1, 128>(persistent, gridCap, params, views, viewCount, tiles, slots, ranges, pairSlots,
                                     proj, lastCount, finalT, queue, stream);
        } else {
            launchBlendViews<1, 256>(persistent, gridCap, params, views, viewCount, tiles, slots, ranges, pairSlots,
                                     proj, lastCount, finalT, queue, stream);
        }
    }
}

void tileOrder(const uint2* ranges, uint32_t tiles, uint32_t* order, hipStream_t stream) {
    if (tiles == 0) return;
    tileOrderKernel<<<1, kOrderWaves * 64, 0, stream>>>(ranges, tiles, order);
}

void blend(const FrameParams& f, uint32_t list, uint32_t slots, const uint32_t* order, const uint2* ranges,
           const uint32_t* pairSlots, const float4* proj, float4* colour, float* depth, uint32_t* lastCount,
           float* finalT, float4* accum, hipStream_t stream) {
    if (f.tilesX * f.tilesY == 0) return;
    if (f.depthMode == 0) {
        launchBlend<0, false>(list, f, slots, order, ranges, pairSlots, proj, nullptr, nullptr, colour, depth,
                              lastCount, finalT, accum, stream);
    } else {
        launchBlend<1, false>(list, f, slots, order, ranges, pairSlots, proj, nullptr, nullptr, colour, depth,
                              lastCount, finalT, accum, stream);
    }
}

void blendComposite(const FrameParams& f, uint32_t list, uint32_t slots, const uint32_t* order, const uint2* ranges,
                    const uint32_t* pairSlots, const float4* proj, const float4* underColour,
                    const float* underDepth, float4* colour, float* depth, hipStream_t stream) {
    if (f.tilesX * f.tilesY == 0) return;
    if (f.depthMode == 0) {
        launchBlend<0, true>(list, f, slots, order, ranges, pairSlots, proj, underColour, underDepth, colour, depth,
                             nullptr, nullptr, nullptr, stream);
    } else {
        launchBlend<1, true>(list, f, slots, order, ranges, pairSlots, proj, underColour, underDepth, colour, depth,
                             nullptr, nullptr, nullptr, stream);
    }
}

}   // namespace lrt::kern
