// Duplication: one (tile, depth) key per tile a splat's footprint touches,
// written at the splat's offset from the exclusive scan of the touched counts.
// Repeats the projection's exact tile test, as the reference's splatEmit does
// (shaders/lrt/splat/splat_emit.slang:17-50), so each splat writes exactly the
// count it was given.
//
// Keys are 64-bit: tile id in the high word, the view depth's float bits in
// the low word (positive floats order as their bits). Values are slots. The
// pairs leave this kernel in slot order, which is what makes one stable sort on
// (tile, depth >> 7) give each tile its splats in the reference's order:
// depth key first, slot on ties.
//
// For a render whose gradients will be asked for, the value is instead the
// pair's own index in this slot-major order (and its slot goes to pairSlot):
// after the sort each pair then knows where its splat's run of pairs lies, so
// per-pair gradients written there sum per splat over contiguous runs.
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "splat_math.h"

namespace lrt::kern {
namespace {

using namespace lrt::dev;

constexpr uint32_t kBlock = 256;

// Writes pair `at` of slot s, or counts it as past the capacity.
template <bool EMIT_INDEX>
__device__ __forceinline__ void emitPair(uint32_t at, uint64_t key, uint32_t s, uint64_t* __restrict__ keys,
                                         uint32_t* __restrict__ values, uint32_t* __restrict__ pairSlot,
                                         uint32_t capacity, uint32_t* __restrict__ overflow) {
    if (at < capacity) {
        keys[at] = key;
        if constexpr (EMIT_INDEX) {
            values[at] = at;
            pairSlot[at] = s;
        } else {
            values[at] = s;
        }
    } else {
        atomicAdd(overflow, 1u);
    }
}

// The wave's slots' pairs: keys (tileBase + tile) << 32 | bits(view z) in the
// frame of width x height with tilesX tiles a row, each slot's in the
// row-major order of its tile rectangle, stopping at the count the projection
// gave. Every lane's rectangle is one run of a wave-wide candidate list
// (countTouchedTilesWave); a round tests 64 consecutive candidates, and a
// touched tile's pair goes to its slot's offset plus the slot's pairs so far
// plus its rank among this round's touched tiles of the same slot -- the place
// the one-lane row-major walk gives it. Every lane of the wave must take part
// (a lane with nothing to write passes expected 0). PER_VIEW: the frame sizes
// differ across lanes and travel with each candidate. With depthRange
// non-null, a slot whose depth key (bits(z) >> 7) lies outside
// [depthRange[0], depthRange[1]] is counted in *outOfRange.
template <bool EMIT_INDEX, bool PER_VIEW>
__device__ __forceinline__ void duplicateWave(uint32_t s, uint32_t expected, uint32_t width, uint32_t height,
                                              uint32_t tilesX, uint32_t tileShift, uint32_t tileBase,
                                              const float4* __restrict__ proj,
                                              const uint2* __restrict__ rects, const uint32_t* __restrict__ offsets,
                                              uint64_t* __restrict__ keys, uint32_t* __restrict__ values,
                                              uint32_t* __restrict__ pairSlot, uint32_t capacity,
                                              uint32_t* __restrict__ overflow,
                                              const uint32_t* __restrict__ depthRange,
                                              uint32_t* __restrict__ outOfRange) {
    const uint32_t lane = __lane_id();
    uint32_t first = 0, x0 = 0, y0 = 0, w = 1, area = 0;
    float4 screen = make_float4(0.0f, 0.0f, 0.0f, 0.0f), conic = screen;
    if (expected != 0) {
        first = offsets[s];
        screen = proj[static_cast<uint64_t>(s) * 3];
        conic = proj[static_cast<uint64_t>(s) * 3 + 1];
        const uint2 rc = rects[s];
        x0 = rc.x & 0xFFFFu, y0 = rc.x >> 16;
        w = (rc.y & 0xFFFFu) - x0 + 1u;
        area = w * ((rc.y >> 16) - y0 + 1u);
    }
    const uint32_t depthBits = floatToBits(screen.z);
    if (depthRange != nullptr && expected != 0) {
        const uint32_t depthKey = (depthBits & 0x7FFFFFFFu) >> 7;
        if (depthKey < depthRange[0] || depthKey > depthRange[1]) atomicAdd(outOfRange, 1u);
    }
    const WaveTiles wt = waveTiles(area);
    const bool exact = wt.total >= (1u << 20);
    const uint64_t below = (1ull << lane) - 1ull;
    const uint32_t slotOfLane0 = s - lane;   // the wave's slots are consecutive
    uint32_t written = 0;                    // this lane's slot's pairs so far
    for (uint32_t base = 0; base < wt.total; base += 64) {
        const uint32_t cand = base + lane;
        const uint32_t o = ownerOf(wt.incl, cand);
        const uint32_t oStart = shflU(wt.start, o), oW = shflU(w, o), ox0 = shflU(x0, o), oy0 = shflU(y0, o);
        const float ocx = shflF(screen.x, o), ocy = shflF(screen.y, o), ocut = shflF(screen.w, o);
        const float oa = shflF(conic.x, o), ob = shflF(conic.y, o), oc = shflF(conic.z, o);
        uint32_t ow = width, oh = height, oTilesX = tilesX, oBase = tileBase;
        if constexpr (PER_VIEW) {
            ow = shflU(width, o);
            oh = shflU(height, o);
            oTilesX = shflU(tilesX, o);
            oBase = shflU(tileBase, o);
        }
        const uint32_t oFirst = shflU(first, o), oExpected = shflU(expected, o), oWritten = shflU(written, o);
        const uint32_t oDepth = shflU(depthBits, o);
        bool hit = false;
        uint32_t tx = 0, ty = 0;
        if (cand < wt.total) {
            uint32_t row, col;
            rowColumn(cand - oStart, oW, exact, row, col);
            tx = ox0 + col;
            ty = oy0 + row;
            hit = tileTouched(ow, oh, ocx, ocy, oa, ob, oc, ocut, tx, ty, tileShift);
        }
        const uint64_t vote = __ballot(hit);
        if (hit) {
            // The owner's run begins at bit runLo of this round (at or below this lane).
            const uint32_t runLo = oStart > base ? oStart - base : 0u;
            const uint64_t sameRunBelow = below & ~((1ull << runLo) - 1ull);
            const uint32_t place = oWritten + static_cast<uint32_t>(__popcll(vote & sameRunBelow));
            if (place < oExpected) {
                const uint64_t key = (static_cast<uint64_t>(oBase + ty * oTilesX + tx) << 32) | oDepth;
                emitPair<EMIT_INDEX>(oFirst + place, key, slotOfLane0 + o, keys, values, pairSlot, capacity, overflow);
            }
        }
        written += static_cast<uint32_t>(__popcll(vote & runMask(wt.start, wt.area, base)));
    }
}

template <bool EMIT_INDEX>
__global__ __launch_bounds__(kBlock) void duplicateKernel(FrameParams f, uint32_t slots,
                                                          const float4* __restrict__ proj,
                                                          const uint2* __restrict__ rects,
                                                          const uint32_t* __restrict__ touched,
                                                          const uint32_t* __restrict__ offsets,
                                                          uint64_t* __restrict__ keys,
                                                          uint32_t* __restrict__ values,
                                                          uint32_t* __restrict__ pairSlot, uint32_t capacity,
                                                          uint32_t* __restrict__ overflow,
                                                          const uint32_t* __restrict__ depthRange,
                                                          uint32_t* __restrict__ outOfRange) {
    const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t expected = s < slots ? touched[s] : 0u;   // every lane stays for the wave's walk
    duplicateWave<EMIT_INDEX, false>(s, expected, f.width, f.height, f.tilesX, f.tileShift, 0u, proj, rects, offsets,
                                     keys, values, pairSlot, capacity, overflow, depthRange, outOfRange);
}

// A batch of views: slot s belongs to view s / splatsPerView.
template <bool EMIT_INDEX>
__global__ __launch_bounds__(kBlock) void duplicateViewsKernel(const ViewInfo* __restrict__ views,
                                                               uint32_t splatsPerView, uint32_t tileShift, uint32_t slots,
                                                               const float4* __restrict__ proj,
                                                               const uint2* __restrict__ rects,
                                                               const uint32_t* __restrict__ touched,
                                                               const uint32_t* __restrict__ offsets,
                                                               uint64_t* __restrict__ keys,
                                                               uint32_t* __restrict__ values,
                                                               uint32_t* __restrict__ pairSlot, uint32_t capacity,
                                                               uint32_t* __restrict__ overflow,
                                                               uint32_t* __restrict__ outOfRange) {
    const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t expected = s < slots ? touched[s] : 0u;   // every lane stays for the wave's walk
    uint32_t width = 0, height = 0, tilesX = 0, tileBase = 0;
    const uint32_t* depthRange = nullptr;
    if (expected != 0) {
        const ViewInfo& v = views[s / splatsPerView];
        width = v.width;
        height = v.height;
        tilesX = v.tilesX;
        tileBase = v.tileBase;
        if (outOfRange != nullptr) depthRange = &v.depthLo;   // depthLo, depthHi
    }
    duplicateWave<EMIT_INDEX, true>(s, expected, width, height, tilesX, tileShift, tileBase, proj, rects, offsets, keys,
                                    values, pairSlot, capacity, overflow, depthRange, outOfRange);
}

}   // namespace

void duplicate(const FrameParams& f, uint32_t slots, const float4* proj, const uint2* rects,
               const uint32_t* touched, const uint32_t* offsets, uint64_t* keys, uint32_t* values,
               uint32_t* pairSlot, uint32_t capacity, uint32_t* overflow, hipStream_t stream,
               const uint32_t* depthRange, uint32_t* outOfRange) {
    if (slots == 0) return;
    const uint32_t blocks = (slots + kBlock - 1) / kBlock;
    if (pairSlot != nullptr) {
        duplicateKernel<true><<<blocks, kBlock, 0, stream>>>(f, slots, proj, rects, touched, offsets, keys, values,
                                                             pairSlot, capacity, overflow, depthRange, outOfRange);
    } else {
        duplicateKernel<false><<<blocks, kBlock, 0, stream>>>(f, slots, proj, rects, touched, offsets, keys, values,
                                                              nullptr, capacity, overflow, depthRange, outOfRange);
    }
}

void duplicateViews(const ViewInfo* views, uint32_t splatsPerView, uint32_t tileShift, uint32_t slots,
                    const float4* proj,
                    const uint2* rects, const uint32_t* touched, const uint32_t* offsets, uint64_t* keys,
                    uint32_t* values, uint32_t* pairSlot, uint32_t capacity, uint32_t* overflow,
                    uint32_t* outOfRange, hipStream_t stream) {
    if (slots == 0 || splatsPerView == 0) return;
    const uint32_t blocks = (slots + kBlock - 1) / kBlock;
    if (pairSlot != nullptr) {
        duplicateViewsKernel<true><<<blocks, kBlock, 0, stream>>>(views, splatsPerView, tileShift, slots, proj, rects,
                                                                  touched,
                                                                  offsets, keys, values, pairSlot, capacity, overflow,
                                                                  outOfRange);
    } else {
        duplicateViewsKernel<false><<<blocks, kBlock, 0, stream>>>(views, splatsPerView, tileShift, slots, proj, rects,
                                                                   touched,
                                                                   offsets, keys, values, nullptr, capacity, overflow,
                                                                   outOfRange);
    }
}

}   // namespace lrt::kern
