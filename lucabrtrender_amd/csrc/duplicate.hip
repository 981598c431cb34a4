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

// Slot s's pairs: keys (tileBase + tile) << 32 | bits(view z) in the frame of
// width x height with tilesX tiles a row, in the row-major order of its tile
// rectangle, stopping at the count the projection gave. A rectangle of at most
// kWaveFootprint tiles is walked by its own lane; larger ones by every active
// lane together, one slot at a time, each pair's place the slot's pairs so far
// plus its rank among this round's touched tiles. Must be reached by every
// active lane (a lane with nothing to write passes touched 0).
template <bool EMIT_INDEX>
__device__ __forceinline__ void duplicateSlot(uint32_t s, uint32_t expected, uint32_t width, uint32_t height,
                                              uint32_t tilesX, uint32_t tileBase, const float4* __restrict__ proj,
                                              const uint2* __restrict__ rects, const uint32_t* __restrict__ offsets,
                                              uint64_t* __restrict__ keys, uint32_t* __restrict__ values,
                                              uint32_t* __restrict__ pairSlot, uint32_t capacity,
                                              uint32_t* __restrict__ overflow) {
    uint32_t first = 0, x0 = 1, y0 = 1, x1 = 0, y1 = 0;
    float4 screen = make_float4(0.0f, 0.0f, 0.0f, 0.0f), conic = screen;
    if (expected != 0) {
        first = offsets[s];
        screen = proj[static_cast<uint64_t>(s) * 3];
        conic = proj[static_cast<uint64_t>(s) * 3 + 1];
        const uint2 rc = rects[s];
        x0 = rc.x & 0xFFFFu, y0 = rc.x >> 16, x1 = rc.y & 0xFFFFu, y1 = rc.y >> 16;
    }
    const uint32_t w = x1 - x0 + 1u;
    const uint32_t area = expected != 0 ? w * (y1 - y0 + 1u) : 0u;
    const bool big = area > kWaveFootprint;
    const uint32_t depthBits = floatToBits(screen.z);
    if (expected != 0 && !big) {
        uint32_t written = 0;
        for (uint32_t ty = y0; ty <= y1 && written < expected; ++ty) {
            for (uint32_t tx = x0; tx <= x1 && written < expected; ++tx) {
                if (!tileTouched(width, height, screen.x, screen.y, conic.x, conic.y, conic.z, screen.w, tx, ty)) {
                    continue;
                }
                const uint64_t key = (static_cast<uint64_t>(tileBase + ty * tilesX + tx) << 32) | depthBits;
                emitPair<EMIT_INDEX>(first + written, key, s, keys, values, pairSlot, capacity, overflow);
                written += 1;
            }
        }
    }
    const ActiveLanes act = activeLanes();
    const uint64_t below = (1ull << act.lane) - 1ull;
    uint64_t todo = __ballot(big);
    while (todo != 0) {
        const uint32_t leader = static_cast<uint32_t>(__builtin_ctzll(todo));
        todo &= todo - 1ull;
        const uint32_t ls = shflU(s, leader), lfirst = shflU(first, leader), lexpected = shflU(expected, leader);
        const uint32_t lw = shflU(w, leader), larea = shflU(area, leader);
        const uint32_t lx0 = shflU(x0, leader), ly0 = shflU(y0, leader);
        const uint32_t lwidth = shflU(width, leader), lheight = shflU(height, leader);
        const uint32_t ltilesX = shflU(tilesX, leader), lbase = shflU(tileBase, leader);
        const uint32_t ldepth = shflU(depthBits, leader);
        const float lcx = shflF(screen.x, leader), lcy = shflF(screen.y, leader), lcut = shflF(screen.w, leader);
        const float la = shflF(conic.x, leader), lb = shflF(conic.y, leader), lc = shflF(conic.z, leader);
        uint32_t written = 0;
        for (uint32_t k0 = 0; k0 < larea && written < lexpected; k0 += act.count) {
            const uint32_t k = k0 + act.rank;
            const uint32_t row = k / lw;
            const uint32_t tx = lx0 + (k - row * lw), ty = ly0 + row;
            const bool hit = k < larea && tileTouched(lwidth, lheight, lcx, lcy, la, lb, lc, lcut, tx, ty);
            const uint64_t vote = __ballot(hit);
            const uint32_t place = written + static_cast<uint32_t>(__popcll(vote & below));
            if (hit && place < lexpected) {
                const uint64_t key = (static_cast<uint64_t>(lbase + ty * ltilesX + tx) << 32) | ldepth;
                emitPair<EMIT_INDEX>(lfirst + place, key, ls, keys, values, pairSlot, capacity, overflow);
            }
            written += static_cast<uint32_t>(__popcll(vote));
        }
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
                                                          uint32_t* __restrict__ overflow) {
    const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= slots) return;
    duplicateSlot<EMIT_INDEX>(s, touched[s], f.width, f.height, f.tilesX, 0u, proj, rects, offsets, keys, values,
                              pairSlot, capacity, overflow);
}

// A batch of views: slot s belongs to view s / splatsPerView.
__global__ __launch_bounds__(kBlock) void duplicateViewsKernel(const ViewInfo* __restrict__ views,
                                                               uint32_t splatsPerView, uint32_t slots,
                                                               const float4* __restrict__ proj,
                                                               const uint2* __restrict__ rects,
                                                               const uint32_t* __restrict__ touched,
                                                               const uint32_t* __restrict__ offsets,
                                                               uint64_t* __restrict__ keys,
                                                               uint32_t* __restrict__ values, uint32_t capacity,
                                                               uint32_t* __restrict__ overflow) {
    const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= slots) return;
    const uint32_t expected = touched[s];
    uint32_t width = 0, height = 0, tilesX = 0, tileBase = 0;
    if (expected != 0) {
        const ViewInfo& v = views[s / splatsPerView];
        width = v.width;
        height = v.height;
        tilesX = v.tilesX;
        tileBase = v.tileBase;
    }
    duplicateSlot<false>(s, expected, width, height, tilesX, tileBase, proj, rects, offsets, keys, values, nullptr,
                         capacity, overflow);
}

}   // namespace

void duplicate(const FrameParams& f, uint32_t slots, const float4* proj, const uint2* rects,
               const uint32_t* touched, const uint32_t* offsets, uint64_t* keys, uint32_t* values,
               uint32_t* pairSlot, uint32_t capacity, uint32_t* overflow, hipStream_t stream) {
    if (slots == 0) return;
    const uint32_t blocks = (slots + kBlock - 1) / kBlock;
    if (pairSlot != nullptr) {
        duplicateKernel<true><<<blocks, kBlock, 0, stream>>>(f, slots, proj, rects, touched, offsets, keys, values,
                                                             pairSlot, capacity, overflow);
    } else {
        duplicateKernel<false><<<blocks, kBlock, 0, stream>>>(f, slots, proj, rects, touched, offsets, keys, values,
                                                              nullptr, capacity, overflow);
    }
}

void duplicateViews(const ViewInfo* views, uint32_t splatsPerView, uint32_t slots, const float4* proj,
                    const uint2* rects, const uint32_t* touched, const uint32_t* offsets, uint64_t* keys,
                    uint32_t* values, uint32_t capacity, uint32_t* overflow, hipStream_t stream) {
    if (slots == 0 || splatsPerView == 0) return;
    duplicateViewsKernel<<<(slots + kBlock - 1) / kBlock, kBlock, 0, stream>>>(
        views, splatsPerView, slots, proj, rects, touched, offsets, keys, values, capacity, overflow);
}

}   // namespace lrt::kern
