// Blend: one wave64 workgroup per 8x8 quadrant of a tile (a 16x16 tile has
// four, an 8x8 one, a 32x32 sixteen), one lane per pixel. The wave walks its
// tile's sorted pairs front to back in rounds of 128 (64 for 64-record lists):
//   1. each lane takes two pairs of the round (one for 64-record lists; their
//      records gathered one round ahead, so the loads are in flight while the
//      previous round is worked)
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
// do, and a quadrant with few pairs frees its slot for the next. The
// quadrants of one tile are blocks b, b + 8, b + 16, ..., which the dispatcher
// deals to one XCD, so their gathers of the tile's records share that XCD's L2.
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

#include <type_traits>

#include "kernels.h"
#include "splat_math.h"

namespace lrt::kern {
namespace {

using namespace lrt::dev;

constexpr uint32_t kLanes = 64;          // one wave: an 8x8 quadrant
constexpr uint32_t kGroup = 4;           // list records walked per step
constexpr float kExpScale = -0.72134752044448170368f;   // -log2(e) / 2
constexpr float kLog2Floor = -7.99435343685885793775f;   // log2(1 / 255): alpha's floor

// The sRGB decode (reference color.slang:7-10) on the hardware log2 and exp2,
// the division by the coverage a reciprocal: a few ulps, as the walk's own
// exponent (the blend's images match the oracle within the image tolerances,
// not bit for bit; the projection, counts and keys stay exact).
__device__ __forceinline__ float srgbToLinearFast(float c) {
    c = fmaxS(c, 0.0f);
    return c <= 0.04045f ? c * (1.0f / 12.92f)
                         : __builtin_amdgcn_exp2f(2.4f * __builtin_amdgcn_logf((c + 0.055f) * (1.0f / 1.055f)));
}

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
        const float inv = __builtin_amdgcn_rcpf(coverage);
        r = srgbToLinearFast(r * inv) * coverage;
        g = srgbToLinearFast(g * inv) * coverage;
        b = srgbToLinearFast(b * inv) * coverage;
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
        const float inv = __builtin_amdgcn_rcpf(coverage);
        r = srgbToLinearFast(r * inv) * coverage;
        g = srgbToLinearFast(g * inv) * coverage;
        b = srgbToLinearFast(b * inv) * coverage;
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

// A quadrant's LDS list: records that reach it, in pair order, padded to a
// multiple of kGroup with empty records (q far below 0: never pass the test).
// One per kernel, whichever walks it runs.
template <uint32_t LIST>
struct BlendList {
    float4 geom[LIST + kGroup];     // A, B, C, D of the expanded form
    float4 conic[LIST + kGroup];    // E, F less the cutoff, o', view z
    // rgb, then bits of the pair's place in the tile's range + 1; composite:
    // the largest view z of the tile's pairs up to this one, kept or not
    float4 colour[LIST + kGroup];
};

// A pixel's walk so far: premultiplied rgb, transmittance, depth sum, the
// threshold depth (-1: not reached), the place of the last contributor, and
// whether it still walks.
struct WalkState {
    float r, g, b, t, depthSum, depthAt;
    uint32_t last;
    bool active;
};

// What a walk leaves (MODE): kFinish writes the finished pixel; kPartial
// writes the walk's state as it stands at the range's end (a segment of a
// split tile, walked from an empty state) to partialA/partialB[lane];
// kResume starts from *state (its active lanes) and leaves the state there.
constexpr uint32_t kFinish = 0, kPartial = 1, kResume = 2;

// A tile of side 2^tileShift holds 2^quadBits 8x8 quadrants (1, 4 or 16),
// numbered row-major inside it.
__device__ __forceinline__ uint32_t quadBits(uint32_t tileShift) { return 2u * (tileShift - 3u); }

// The first pixel of quadrant `quadrant` of tile `tile` of the frame f.
__device__ __forceinline__ uint2 quadrantOrigin(const FrameParams& f, uint32_t tile, uint32_t quadrant) {
    const uint32_t side = f.tileShift - 3u;
    return make_uint2(((tile % f.tilesX) << f.tileShift) + ((quadrant & ((1u << side) - 1u)) << 3),
                      ((tile / f.tilesX) << f.tileShift) + ((quadrant >> side) << 3));
}

// Quadrant `quadrant` of tile `tile` of the frame f, walking the pairs
// [range.x, range.y) of pairSlots (all of the tile's, or a segment of them;
// places count from tileStart); colour and depth are the frame's images,
// lastCount/finalT/accum its per-pixel state (each may be null).
template <uint32_t DEPTH_MODE, uint32_t LIST, bool COMPOSITE, uint32_t MODE = kFinish, bool TRACK = true>
__device__ __forceinline__ void blendQuadrant(BlendList<LIST>& list, const FrameParams& f, uint32_t tile,
                                              uint32_t quadrant, uint2 range,
                                              uint32_t tileStart, uint32_t slots,
                                              const uint32_t* __restrict__ pairSlots,
                                              const float4* __restrict__ proj,
                                              const float4* __restrict__ underColour,
                                              const float* __restrict__ underDepth, float4* __restrict__ colour,
                                              float* __restrict__ depth, uint32_t* __restrict__ lastCount,
                                              float* __restrict__ finalT, float4* __restrict__ accum,
                                              WalkState* state = nullptr, float4* __restrict__ partialA = nullptr,
                                              float2* __restrict__ partialB = nullptr,
                                              uint32_t* __restrict__ walkSteps = nullptr,
                                              const uint32_t* __restrict__ slotOfPair = nullptr) {
    static_assert(!(COMPOSITE && MODE != kFinish), "a composite walk is not split");
    // Pairs a lane stages a round: two, or one for a 64-record list.
    constexpr uint32_t kPerLane = LIST >= 2 * kLanes ? 2u : 1u;
    constexpr uint32_t kRound = kLanes * kPerLane;
    static_assert(LIST % kRound == 0 && LIST >= kRound, "list length");
    float4* const sGeom = list.geom;
    float4* const sConic = list.conic;
    float4* const sColour = list.colour;

    const uint32_t lane = threadIdx.x;
    const uint2 origin = quadrantOrigin(f, tile, quadrant);
    const uint32_t x0 = origin.x, y0 = origin.y;
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
    // A training frame's sort carried pair indices: each pair's slot is read through it.
    auto slotAt = [&](uint32_t at) -> uint32_t {
        if (at >= range.y) return 0xFFFFFFFFu;
        const uint32_t v = pairSlots[at];
        return slotOfPair != nullptr ? slotOfPair[v] : v;
    };
    Staged staged[kPerLane];
#pragma unroll
    for (uint32_t u = 0; u < kPerLane; ++u) staged[u] = gather(proj, slotAt(range.x + u * kLanes + lane), slots);

    float r = 0.0f, g = 0.0f, b = 0.0f, t = 1.0f, depthSum = 0.0f, depthAt = -1.0f;
    uint32_t last = 0;
    bool active = inside;
    if constexpr (MODE == kResume) {
        r = state->r;
        g = state->g;
        b = state->b;
        t = state->t;
        depthSum = state->depthSum;
        depthAt = state->depthAt;
        last = state->last;
        active = inside && state->active;
    }
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
    uint32_t walked = 0;       // list records walked (kGroup a step)
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
                    const float fbase = ux * (fa * ux + fb * vy) + fc * vy * vy;
                    sGeom[to] = make_float4(fa, fb, fc, fd);
                    if constexpr (COMPOSITE) {
                        // q = the form less the cutoff; alpha = o' 2^q. A disc
                        // (colour.w 1) is marked by o' < 0: its alpha never passes.
                        const bool disc = s.col.w > 0.5f;
                        const float scaledOpacity = disc ? -1.0f : s.conic.w * __builtin_amdgcn_exp2f(scaledCut);
                        sConic[to] = make_float4(fe, fbase - scaledCut, scaledOpacity, s.screen.z);
                    } else {
                        // q' = the form plus log2(opacity): alpha = 2^q' directly, and
                        // both tests (inside the cutoff, alpha >= 1/255) are q' >= qmin
                        const float logOpacity = __builtin_amdgcn_logf(s.conic.w);
                        const float qmin = fmaxS(logOpacity + scaledCut, kLog2Floor);
                        sConic[to] = make_float4(fe, fbase + logOpacity, qmin, s.screen.z);
                    }
                    sColour[to] = make_float4(s.col.x, s.col.y, s.col.z,
                                              COMPOSITE ? zReach : __uint_as_float(at - tileStart + 1u));
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
            sConic[fill + lane] = make_float4(0.0f, -1e30f, 0.0f, 0.0f);   // q far below any threshold: never passes
            sColour[fill + lane] = make_float4(0.0f, 0.0f, 0.0f, COMPOSITE ? -3.4e38f : 0.0f);
        }
        __syncthreads();
        // 3. walk the list: all kGroup records' reads issued together, then the
        // pixels blend them in order
        for (uint32_t j = 0; j < padded; j += kGroup) {
            walked += kGroup;
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
                float alpha;
                bool ok;   // bitwise, not short-circuit: no branch inside the walk
                if constexpr (COMPOSITE) {
                    alpha = __builtin_fminf(0.99f, g1[k].z * __builtin_amdgcn_exp2f(q));
                    ok = active & (q >= 0.0f) & (alpha >= 1.0f / 255.0f);
                } else {
                    alpha = __builtin_fminf(0.99f, __builtin_amdgcn_exp2f(q));
                    ok = active & (q >= g1[k].z);
                }
                const float a = ok ? alpha : 0.0f;
                const float weight = t * a;
                r = __builtin_fmaf(weight, g2[k].x, r);
                g = __builtin_fmaf(weight, g2[k].y, g);
                b = __builtin_fmaf(weight, g2[k].z, b);
                depthSum = __builtin_fmaf(weight, g1[k].w, depthSum);
                t = t - weight;   // T (1 - alpha)
                if constexpr (!COMPOSITE && TRACK) last = ok ? __float_as_uint(g2[k].w) : last;
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
    // The walk's records, counted on request (the blend's VALU work is
    // 64 lanes x these steps).
    if (walkSteps != nullptr && lane == 0) atomicAdd(walkSteps, walked);
    if constexpr (MODE == kPartial) {
        partialA[lane] = make_float4(r, g, b, depthSum);
        partialB[lane] = make_float2(t, __uint_as_float(last));
        return;
    } else if constexpr (MODE == kResume) {
        *state = WalkState{r, g, b, t, depthSum, depthAt, last, active};
        return;
    }
    if (!inside) return;
    if constexpr (COMPOSITE) {
        finishComposite(f, at0, r, g, b, t, depthSum, depthAt, opaqueHit, make_float4(oR, oG, oB, 1.0f), oZ,
                        underHit, underColour, underZ, colour, depth);
    } else {
        finishPixel(f, px, py, r, g, b, t, depthSum, depthAt, last, colour, depth, lastCount, finalT, accum);
    }
}

template <uint32_t DEPTH_MODE, uint32_t LIST, bool COMPOSITE, bool TRACK>
__global__ __launch_bounds__(kLanes) void blendKernel(FrameParams f, uint32_t slots, const uint32_t* __restrict__ order,
                                                      const uint2* __restrict__ ranges,
                                                      const uint32_t* __restrict__ pairSlots,
                                                      const float4* __restrict__ proj,
                                                      const float4* __restrict__ underColour,
                                                      const float* __restrict__ underDepth,
                                                      float4* __restrict__ colour, float* __restrict__ depth,
                                                      uint32_t* __restrict__ lastCount,
                                                      float* __restrict__ finalT, float4* __restrict__ accum,
                                                      uint32_t* __restrict__ walkSteps,
                                                      const uint32_t* __restrict__ slotOfPair) {
    const uint32_t tiles = f.tilesX * f.tilesY;
    const uint32_t qBits = quadBits(f.tileShift);
    const uint32_t quadrant = (blockIdx.x >> 3) & ((1u << qBits) - 1u);
    const uint32_t place = (blockIdx.x >> (3u + qBits)) * 8u + (blockIdx.x & 7u);
    if (place >= tiles) return;
    const uint32_t tile = order != nullptr ? order[place] : place;
    if (tile >= tiles) return;   // an order entry that is not a tile: draw nothing
    __shared__ BlendList<LIST> list;
    const uint2 rg = ranges[tile];
    blendQuadrant<DEPTH_MODE, LIST, COMPOSITE, kFinish, TRACK>(list, f, tile, quadrant, rg, rg.x, slots, pairSlots, proj,
                                                               underColour, underDepth, colour, depth, lastCount,
                                                               finalT, accum, nullptr, nullptr, nullptr, walkSteps,
                                                               slotOfPair);
}

// Every quadrant of every view of a batch (2^qBits quadrants a tile). As a
// grid (PERSISTENT false) block b is task b, the quadrants of a tile going to
// blocks b, b + 8, b + 16, ... of one XCD; persistent, each block takes tasks
// from a shared counter until none are left (task t: tile t >> qBits,
// quadrant t & (2^qBits - 1)), so a block that drew a light quadrant takes the
// next instead of leaving its slot idle.
template <uint32_t DEPTH_MODE, uint32_t LIST, bool PERSISTENT, bool TRACK>
__global__ __launch_bounds__(kLanes) void blendViewsKernel(const FrameParams* __restrict__ params,
                                                           const ViewInfo* __restrict__ views, uint32_t viewCount,
                                                           uint32_t tiles, uint32_t qBits, uint32_t tasks,
                                                           uint32_t slots,
                                                           const uint2* __restrict__ ranges,
                                                           const uint32_t* __restrict__ pairSlots,
                                                           const float4* __restrict__ proj,
                                                           uint32_t* __restrict__ lastCount,
                                                           float* __restrict__ finalT, float4* __restrict__ accum,
                                                           uint32_t* __restrict__ queue,
                                                           const uint32_t* __restrict__ slotOfPair) {
    __shared__ BlendList<LIST> list;
    for (uint32_t round = 0;; ++round) {
        uint32_t task, g, quadrant;
        if constexpr (PERSISTENT) {
            uint32_t taken = 0;
            if (threadIdx.x == 0) taken = atomicAdd(queue, 1u);
            task = static_cast<uint32_t>(__shfl(static_cast<int>(taken), 0, 64));
            if (task >= tasks) return;
            g = task >> qBits;
            quadrant = task & ((1u << qBits) - 1u);
            if (round > 0) __syncthreads();   // the last task's LDS list is done with
        } else {
            task = blockIdx.x;
            g = (task >> (3u + qBits)) * 8u + (task & 7u);
            quadrant = (task >> 3) & ((1u << qBits) - 1u);
        }
        if (g < tiles) {
            const ViewInfo& v = views[viewOfTile(views, viewCount, g)];
            const uint2 rg = ranges[g];
            blendQuadrant<DEPTH_MODE, LIST, false, kFinish, TRACK>(
                list, params[v.params], g - v.tileBase, quadrant, rg, rg.x, slots, pairSlots, proj, nullptr, nullptr,
                reinterpret_cast<float4*>(v.colour), v.depth, lastCount != nullptr ? lastCount + v.pixelBase : nullptr,
                finalT != nullptr ? finalT + v.pixelBase : nullptr, accum != nullptr ? accum + v.pixelBase : nullptr,
                nullptr, nullptr, nullptr, nullptr, slotOfPair);
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
constexpr uint32_t kNoSlot = 0xFFFFFFFFu;   // a task that finishes its tile's pixels itself
constexpr uint32_t kOrderWaves = 16;
constexpr uint32_t kOrderCached = 32768;   // tiles whose bucket the first walk leaves in LDS for the second

__device__ __forceinline__ uint32_t bucketOf(uint2 rg) {
    return 32u - static_cast<uint32_t>(__clz(rg.y - rg.x));   // 0 for an empty tile
}

// The segments of a tile of `pairs` pairs: one, unless it holds more than
// segmentPairs (0: never split).
__device__ __forceinline__ uint32_t segmentsOf(uint32_t pairs, uint32_t segmentPairs) {
    return segmentPairs == 0 || pairs <= segmentPairs ? 1u : (pairs + segmentPairs - 1u) / segmentPairs;
}

// Exclusive sums over the workgroup of three counts, and their totals.
__device__ __forceinline__ void blockScan3(uint32_t v[3], uint32_t total[3], uint32_t (*sWave)[kOrderWaves]) {
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint32_t inc[3] = {v[0], v[1], v[2]};
#pragma unroll
    for (uint32_t d = 1; d < 64; d <<= 1) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const uint32_t o = static_cast<uint32_t>(__shfl_up(static_cast<int>(inc[k]), d, 64));
            inc[k] += lane >= d ? o : 0u;
        }
    }
    if (lane == 63)
        for (int k = 0; k < 3; ++k) sWave[k][wave] = inc[k];
    __syncthreads();
    for (int k = 0; k < 3; ++k) {
        uint32_t before = 0, all = 0;
        for (uint32_t w = 0; w < kOrderWaves; ++w) {
            before += w < wave ? sWave[k][w] : 0u;
            all += sWave[k][w];
        }
        total[k] = all;
        v[k] = before + inc[k] - v[k];
    }
}

__global__ __launch_bounds__(kOrderWaves * 64) void tileOrderKernel(const uint2* __restrict__ ranges, uint32_t tiles,
                                                                    uint32_t* __restrict__ order,
                                                                    uint32_t segmentPairs, uint4* __restrict__ tasks,
                                                                    uint4* __restrict__ combine,
                                                                    uint32_t* __restrict__ planCounts) {
    __shared__ uint32_t sCount[kOrderWaves][kBuckets];
    __shared__ uint32_t sStart[kOrderWaves][kBuckets];
    __shared__ uint8_t sBucket[kOrderCached];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    for (uint32_t k = tid; k < kOrderWaves * kBuckets; k += blockDim.x) (&sCount[0][0])[k] = 0;
    __syncthreads();
    const uint32_t segment = (tiles + kOrderWaves * 64 - 1) / (kOrderWaves * 64) * 64;   // whole chunks of 64
    const uint32_t begin = wave * segment;
    const uint32_t end = begin + segment < tiles ? begin + segment : tiles;
    // Eight loads in flight before the first count: one trip to memory for a
    // wave's tiles of a 1080p frame instead of one a chunk.
    constexpr uint32_t kOrderBatch = 8;
    for (uint32_t t0 = begin + lane; t0 < end; t0 += 64 * kOrderBatch) {
        uint2 rg[kOrderBatch];
#pragma unroll
        for (uint32_t b = 0; b < kOrderBatch; ++b) {
            const uint32_t t = t0 + 64 * b;
            rg[b] = t < end ? ranges[t] : make_uint2(0u, 0u);
        }
#pragma unroll
        for (uint32_t b = 0; b < kOrderBatch; ++b) {
            const uint32_t t = t0 + 64 * b;
            if (t < end) {
                const uint32_t bucket = bucketOf(rg[b]);
                if (t < kOrderCached) sBucket[t] = static_cast<uint8_t>(bucket);
                atomicAdd(&sCount[wave][bucket], 1u);
            }
        }
    }
    __syncthreads();
    // Bucket k's first place: the tiles of every heavier bucket; then wave w's
    // within it: the bucket's tiles in the waves before w.
    // One wave: lane k < kBuckets holds bucket k's per-wave counts, scans them
    // over the waves in registers, and takes its bucket's first place as the
    // tiles of the heavier buckets (lanes above it: a suffix sum over lanes).
    if (wave == 0) {
        uint32_t c[kOrderWaves], sum = 0;
#pragma unroll
        for (uint32_t ww = 0; ww < kOrderWaves; ++ww) {
            c[ww] = lane < kBuckets ? sCount[ww][lane] : 0u;
            sum += c[ww];
        }
        uint32_t suffix = sum;   // lanes lane .. 63 (zero past the buckets)
#pragma unroll
        for (uint32_t d = 1; d < 64; d <<= 1) {
            const uint32_t o = static_cast<uint32_t>(__shfl_down(static_cast<int>(suffix), d, 64));
            suffix += lane + d < 64 ? o : 0u;
        }
        const uint32_t above = suffix - sum;
        if (lane < kBuckets) {
            uint32_t start = above;
#pragma unroll
            for (uint32_t ww = 0; ww < kOrderWaves; ++ww) {
                sStart[ww][lane] = start;
                start += c[ww];
            }
        }
    }
    __syncthreads();
    // Each lane's rank among the lanes of its bucket (six ballots on the
    // bucket's bits give the lanes that share it); the group's last lane then
    // moves the bucket's next place on. The places live in LDS, read and written
    // through a volatile pointer so that every round sees the last one's.
    volatile uint32_t* next = &sStart[wave][0];
    const uint64_t below = (1ull << lane) - 1ull;
    static_assert(kBuckets < 64, "a bucket (and the no-tile mark kBuckets) fits six bits");
    for (uint32_t base = begin; base < end; base += 64) {
        const uint32_t t = base + lane;
        const bool mine = t < end;
        const uint32_t bucket = !mine ? kBuckets : t < kOrderCached ? sBucket[t] : bucketOf(ranges[t]);
        uint64_t same = ~0ull;
#pragma unroll
        for (uint32_t bit = 0; bit < 6; ++bit) {
            const bool set = ((bucket >> bit) & 1u) != 0u;
            const uint64_t vote = __ballot(set);
            same &= set ? vote : ~vote;
        }
        const uint32_t place = mine ? next[bucket] + static_cast<uint32_t>(__popcll(same & below)) : 0u;
        if (mine) order[place] = t;
        if (mine && (same & ~below & ~(1ull << lane)) == 0) next[bucket] = place + 1u;
    }
    if (tasks == nullptr) return;
    // The split plan: tiles in that order, each cut into segmentsOf() segments
    // of near-equal length. tasks[k] = (tile, segment begin, end, partial slot
    // or kNoSlot); combine[m] = (tile, first slot, segments) for each split
    // tile. Each thread takes a run of places; sums over the runs place them.
    __shared__ uint32_t sWave[3][kOrderWaves];
    __syncthreads();   // the order is written
    const uint32_t chunk = (tiles + blockDim.x - 1u) / blockDim.x;
    const uint32_t p0 = tid * chunk < tiles ? tid * chunk : tiles;
    const uint32_t p1 = p0 + chunk < tiles ? p0 + chunk : tiles;
    uint32_t v[3] = {0u, 0u, 0u};   // tasks, partial slots, split tiles
    for (uint32_t p = p0; p < p1; ++p) {
        const uint2 rg = ranges[order[p]];
        const uint32_t n = segmentsOf(rg.y - rg.x, segmentPairs);
        v[0] += n;
        v[1] += n > 1 ? n : 0u;
        v[2] += n > 1 ? 1u : 0u;
    }
    uint32_t total[3];
    blockScan3(v, total, sWave);
    for (uint32_t p = p0; p < p1; ++p) {
        const uint32_t tile = order[p];
        const uint2 rg = ranges[tile];
        const uint32_t len = rg.y - rg.x;
        const uint32_t n = segmentsOf(len, segmentPairs);
        const uint32_t segLen = (len + n - 1u) / n;
        for (uint32_t k = 0; k < n; ++k) {
            const uint32_t begin = rg.x + k * segLen < rg.y ? rg.x + k * segLen : rg.y;
            const uint32_t end = begin + segLen < rg.y ? begin + segLen : rg.y;
            tasks[v[0] + k] = make_uint4(tile, begin, end, n > 1 ? v[1] + k : kNoSlot);
        }
        if (n > 1) {
            combine[v[2]] = make_uint4(tile, v[1], n, 0u);
            v[1] += n;
            v[2] += 1;
        }
        v[0] += n;
    }
    if (tid == 0) {
        planCounts[0] = total[0];
        planCounts[1] = total[2];
    }
}

// The tasks of a split plan: task k's quadrants are blocks b, b + 8, b + 16,
// ... of one XCD. A task of an unsplit tile finishes its pixels; a
// segment leaves its partial walk in its slot.
template <uint32_t DEPTH_MODE, uint32_t LIST>
__global__ __launch_bounds__(kLanes) void blendTasksKernel(FrameParams f, uint32_t slots, const uint4* __restrict__ tasks,
                                                           const uint32_t* __restrict__ planCounts,
                                                           const uint2* __restrict__ ranges,
                                                           const uint32_t* __restrict__ pairSlots,
                                                           const float4* __restrict__ proj,
                                                           float4* __restrict__ colour, float* __restrict__ depth,
                                                           uint32_t* __restrict__ lastCount,
                                                           float* __restrict__ finalT, float4* __restrict__ partialA,
                                                           float2* __restrict__ partialB) {
    const uint32_t qBits = quadBits(f.tileShift);
    const uint32_t quadrant = (blockIdx.x >> 3) & ((1u << qBits) - 1u);
    const uint32_t task = (blockIdx.x >> (3u + qBits)) * 8u + (blockIdx.x & 7u);
    if (task >= planCounts[0]) return;
    __shared__ BlendList<LIST> list;
    const uint4 tk = tasks[task];
    const uint32_t tileStart = ranges[tk.x].x;
    if (tk.w == kNoSlot) {
        blendQuadrant<DEPTH_MODE, LIST, false, kFinish>(list, f, tk.x, quadrant, make_uint2(tk.y, tk.z), tileStart, slots,
                                                        pairSlots, proj, nullptr, nullptr, colour, depth, lastCount,
                                                        finalT, nullptr);
    } else {
        const uint64_t at = ((static_cast<uint64_t>(tk.w) << qBits) + quadrant) * kLanes;
        blendQuadrant<DEPTH_MODE, LIST, false, kPartial>(list, f, tk.x, quadrant, make_uint2(tk.y, tk.z), tileStart, slots,
                                                         pairSlots, proj, nullptr, nullptr, colour, depth, lastCount,
                                                         finalT, nullptr, nullptr, partialA + at, partialB + at);
    }
}

// A split tile's pixels, its segments' partial walks composited front to back:
// rgb += T rgb_s, depth sum += T depth_s, T *= T_s. A pixel whose walk would
// stop inside a segment (T below 1e-4, or the threshold depth crossed) walks
// that segment again from its true state, so it stops where the one walk
// would; the others take the segment whole.
template <uint32_t DEPTH_MODE, uint32_t LIST>
__global__ __launch_bounds__(kLanes) void combineKernel(FrameParams f, uint32_t slots, const uint4* __restrict__ combine,
                                                        const uint32_t* __restrict__ planCounts,
                                                        const uint2* __restrict__ ranges,
                                                        const uint32_t* __restrict__ pairSlots,
                                                        const float4* __restrict__ proj, float4* __restrict__ colour,
                                                        float* __restrict__ depth, uint32_t* __restrict__ lastCount,
                                                        float* __restrict__ finalT,
                                                        const float4* __restrict__ partialA,
                                                        const float2* __restrict__ partialB) {
    const uint32_t qBits = quadBits(f.tileShift);
    const uint32_t quadrant = blockIdx.x & ((1u << qBits) - 1u);
    const uint32_t c = blockIdx.x >> qBits;
    if (c >= planCounts[1]) return;
    const uint4 ct = combine[c];
    const uint32_t tile = ct.x, n = ct.z, lane = threadIdx.x;
    const uint2 origin = quadrantOrigin(f, tile, quadrant);
    const uint32_t x0 = origin.x, y0 = origin.y;
    if (x0 >= f.width || y0 >= f.height) return;
    const uint32_t px = x0 + (lane & 7u), py = y0 + (lane >> 3);
    const bool inside = px < f.width && py < f.height && px < x0 + 8u && py < y0 + 8u;
    const uint2 rg = ranges[tile];
    const uint32_t len = rg.y - rg.x, segLen = (len + n - 1u) / n;
    __shared__ BlendList<LIST> list;
    WalkState st{0.0f, 0.0f, 0.0f, 1.0f, 0.0f, -1.0f, 0u, inside};
    for (uint32_t k = 0; k < n; ++k) {
        const uint64_t at = ((static_cast<uint64_t>(ct.y + k) << qBits) + quadrant) * kLanes + lane;
        const float4 pa = partialA[at];
        const float2 pb = partialB[at];
        bool again = false;
        if (st.active) {
            const float tEnd = st.t * pb.x;
            again = tEnd < 1e-4f;
            if constexpr (DEPTH_MODE == 1) again = again || (st.depthAt < 0.0f && 1.0f - tEnd >= f.depthThreshold);
            if (!again) {
                st.r = __builtin_fmaf(st.t, pa.x, st.r);
                st.g = __builtin_fmaf(st.t, pa.y, st.g);
                st.b = __builtin_fmaf(st.t, pa.z, st.b);
                st.depthSum = __builtin_fmaf(st.t, pa.w, st.depthSum);
                st.t = tEnd;
                st.last = __float_as_uint(pb.y) != 0u ? __float_as_uint(pb.y) : st.last;
            }
        }
        if (__ballot(again) != 0) {
            const uint32_t begin = rg.x + k * segLen < rg.y ? rg.x + k * segLen : rg.y;
            const uint32_t end = begin + segLen < rg.y ? begin + segLen : rg.y;
            WalkState w = st;
            w.active = again;
            blendQuadrant<DEPTH_MODE, LIST, false, kResume>(list, f, tile, quadrant, make_uint2(begin, end), rg.x, slots,
                                                            pairSlots, proj, nullptr, nullptr, colour, depth,
                                                            lastCount, finalT, nullptr, &w);
            if (again) st = w;
            __syncthreads();   // the walk's list is refilled by the next one
        }
    }
    if (!inside) return;
    finishPixel(f, px, py, st.r, st.g, st.b, st.t, st.depthSum, st.depthAt, st.last, colour, depth, lastCount, finalT,
                nullptr);
}

template <uint32_t DEPTH_MODE, bool COMPOSITE>
void launchBlend(uint32_t list, const FrameParams& f, uint32_t slots, const uint32_t* order, const uint2* ranges,
                 const uint32_t* pairSlots, const float4* proj, const float4* underColour, const float* underDepth,
                 float4* colour, float* depth, uint32_t* lastCount, float* finalT, float4* accum,
                 hipStream_t stream, uint32_t* walkSteps = nullptr, const uint32_t* slotOfPair = nullptr) {
    // A block a quadrant, the grid rounded up to whole groups of eight tiles.
    // The last contributor is tracked only where it is kept.
    const uint32_t tiles = f.tilesX * f.tilesY;
    const uint32_t blocks = (tiles + 7u) / 8u * (8u << (2u * (f.tileShift - 3u)));
    const auto go = [&](auto listLen, auto track) {
        blendKernel<DEPTH_MODE, decltype(listLen)::value, COMPOSITE, decltype(track)::value>
            <<<blocks, kLanes, 0, stream>>>(f, slots, order, ranges, pairSlots, proj, underColour, underDepth, colour,
                                            depth, lastCount, finalT, accum, walkSteps, slotOfPair);
    };
    using L64 = std::integral_constant<uint32_t, 64>;
    using L128 = std::integral_constant<uint32_t, 128>;
    using L256 = std::integral_constant<uint32_t, 256>;
    using Keep = std::true_type;
    using Drop = std::false_type;
    const bool track = !COMPOSITE && lastCount != nullptr;
    if (list <= 64) {
        if (track) go(L64{}, Keep{}); else go(L64{}, Drop{});
    } else if (list <= 128) {
        if (track) go(L128{}, Keep{}); else go(L128{}, Drop{});
    } else {
        if (track) go(L256{}, Keep{}); else go(L256{}, Drop{});
    }
}

template <uint32_t DEPTH_MODE, uint32_t LIST>
void launchBlendViews(bool persistent, uint32_t gridCap, const FrameParams* params, const ViewInfo* views,
                      uint32_t viewCount, uint32_t tiles, uint32_t tileShift, uint32_t slots, const uint2* ranges,
                      const uint32_t* pairSlots, const float4* proj, uint32_t* lastCount, float* finalT, float4* accum,
                      uint32_t* queue, hipStream_t stream, const uint32_t* slotOfPair) {
    const uint32_t qBits = 2u * (tileShift - 3u);
    const auto go = [&](auto track) {
        constexpr bool T = decltype(track)::value;
        if (persistent) {
            const uint32_t tasks = tiles << qBits;
            const uint32_t grid = tasks < gridCap ? tasks : gridCap;
            fill32(queue, 0u, 1, stream);
            blendViewsKernel<DEPTH_MODE, LIST, true, T><<<grid, kLanes, 0, stream>>>(
                params, views, viewCount, tiles, qBits, tasks, slots, ranges, pairSlots, proj, lastCount, finalT,
                accum, queue, slotOfPair);
        } else {
            const uint32_t tasks = (tiles + 7u) / 8u * (8u << qBits);
            blendViewsKernel<DEPTH_MODE, LIST, false, T><<<tasks, kLanes, 0, stream>>>(
                params, views, viewCount, tiles, qBits, tasks, slots, ranges, pairSlots, proj, lastCount, finalT,
                accum, queue, slotOfPair);
        }
    };
    if (lastCount != nullptr) go(std::true_type{}); else go(std::false_type{});
}

}   // namespace

void blendViews(uint32_t depthMode, uint32_t list, bool persistent, uint32_t gridCap, const FrameParams* params,
                const ViewInfo* views, uint32_t viewCount, uint32_t tiles, uint32_t tileShift, uint32_t slots,
                const uint2* ranges,
                const uint32_t* pairSlots, const float4* proj, uint32_t* lastCount, float* finalT, float4* accum,
                uint32_t* queue, hipStream_t stream, const uint32_t* slotOfPair) {
    if (tiles == 0 || viewCount == 0) return;
    const auto go = [&](auto depthMode, auto listLen) {
        launchBlendViews<decltype(depthMode)::value, decltype(listLen)::value>(
            persistent, gridCap, params, views, viewCount, tiles, tileShift, slots, ranges, pairSlots, proj, lastCount,
            finalT, accum, queue, stream, slotOfPair);
    };
    using D0 = std::integral_constant<uint32_t, 0>;
    using D1 = std::integral_constant<uint32_t, 1>;
    using L64 = std::integral_constant<uint32_t, 64>;
    using L128 = std::integral_constant<uint32_t, 128>;
    using L256 = std::integral_constant<uint32_t, 256>;
    if (depthMode == 0) {
        if (list <= 64) go(D0{}, L64{}); else if (list <= 128) go(D0{}, L128{}); else go(D0{}, L256{});
    } else {
        if (list <= 64) go(D1{}, L64{}); else if (list <= 128) go(D1{}, L128{}); else go(D1{}, L256{});
    }
}

void tileOrder(const uint2* ranges, uint32_t tiles, uint32_t* order, hipStream_t stream) {
    if (tiles == 0) return;
    tileOrderKernel<<<1, kOrderWaves * 64, 0, stream>>>(ranges, tiles, order, 0u, nullptr, nullptr, nullptr);
}

void blendSplit(const FrameParams& f, uint32_t list, uint32_t segmentPairs, uint32_t slots, const uint2* ranges,
                const uint32_t* pairSlots, const float4* proj, const BlendPlan& plan, float4* colour, float* depth,
                uint32_t* lastCount, float* finalT, hipStream_t stream) {
    const uint32_t tiles = f.tilesX * f.tilesY;
    if (tiles == 0) return;
    tileOrderKernel<<<1, kOrderWaves * 64, 0, stream>>>(ranges, tiles, plan.order, segmentPairs, plan.tasks,
                                                        plan.combine, plan.counts);
    const uint32_t qBits = 2u * (f.tileShift - 3u);
    const uint32_t taskBlocks = (plan.taskCapacity + 7u) / 8u * (8u << qBits);
    const uint32_t combineBlocks = plan.combineCapacity << qBits;
    const auto launch = [&](auto depthMode, auto listLen) {
        constexpr uint32_t D = decltype(depthMode)::value, L = decltype(listLen)::value;
        blendTasksKernel<D, L><<<taskBlocks, kLanes, 0, stream>>>(f, slots, plan.tasks, plan.counts, ranges, pairSlots,
                                                                  proj, colour, depth, lastCount, finalT,
                                                                  plan.partialA, plan.partialB);
        if (combineBlocks > 0) {
            combineKernel<D, L><<<combineBlocks, kLanes, 0, stream>>>(f, slots, plan.combine, plan.counts, ranges,
                                                                      pairSlots, proj, colour, depth, lastCount,
                                                                      finalT, plan.partialA, plan.partialB);
        }
    };
    using D0 = std::integral_constant<uint32_t, 0>;
    using D1 = std::integral_constant<uint32_t, 1>;
    using L64 = std::integral_constant<uint32_t, 64>;
    using L128 = std::integral_constant<uint32_t, 128>;
    using L256 = std::integral_constant<uint32_t, 256>;
    if (f.depthMode == 0) {
        if (list <= 64) launch(D0{}, L64{}); else if (list <= 128) launch(D0{}, L128{}); else launch(D0{}, L256{});
    } else {
        if (list <= 64) launch(D1{}, L64{}); else if (list <= 128) launch(D1{}, L128{}); else launch(D1{}, L256{});
    }
}

void blend(const FrameParams& f, uint32_t list, uint32_t slots, const uint32_t* order, const uint2* ranges,
           const uint32_t* pairSlots, const float4* proj, float4* colour, float* depth, uint32_t* lastCount,
           float* finalT, float4* accum, hipStream_t stream, uint32_t* walkSteps, const uint32_t* slotOfPair) {
    if (f.tilesX * f.tilesY == 0) return;
    if (f.depthMode == 0) {
        launchBlend<0, false>(list, f, slots, order, ranges, pairSlots, proj, nullptr, nullptr, colour, depth,
                              lastCount, finalT, accum, stream, walkSteps, slotOfPair);
    } else {
        launchBlend<1, false>(list, f, slots, order, ranges, pairSlots, proj, nullptr, nullptr, colour, depth,
                              lastCount, finalT, accum, stream, walkSteps, slotOfPair);
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
