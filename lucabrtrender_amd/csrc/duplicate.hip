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

// Slot s's pairs: keys (tileBase + tile) << 32 | bits(view z) in the frame of
// width x height with tilesX tiles a row.
template <bool EMIT_INDEX>
__device__ __forceinline__ void duplicateSlot(uint32_t s, uint32_t width, uint32_t height, uint32_t tilesX,
                                              uint32_t tileBase, const float4* __restrict__ proj,
                                              const uint2* __restrict__ rects, const uint32_t* __restrict__ touched,
                                              const uint32_t* __restrict__ offsets, uint64_t* __restrict__ keys,
                                              uint32_t* __restrict__ values, uint32_t* __restrict__ pairSlot,
                                              uint32_t capacity, uint32_t* __restrict__ overflow) {
    const uint32_t expected = touched[s];
    if (expected == 0) return;
    const uint32_t first = offsets[s];
    const float4 screen = proj[static_cast<uint64_t>(s) * 3];
    const float4 conic = proj[static_cast<uint64_t>(s) * 3 + 1];
    const uint2 rc = rects[s];
    const uint32_t x0 = rc.x & 0xFFFFu, y0 = rc.x >> 16, x1 = rc.y & 0xFFFFu, y1 = rc.y >> 16;
    const uint64_t depthBits = floatToBits(screen.z);
    uint32_t written = 0;
    for (uint32_t ty = y0; ty <= y1 && written < expected; ++ty) {
        for (uint32_t tx = x0; tx <= x1 && written < expected; ++tx) {
            if (!tileTouched(width, height, screen.x, screen.y, conic.x, conic.y, conic.z, screen.w, tx, ty)) {
                continue;
            }
            const uint32_t at = first + written;
            if (at < capacity) {
                keys[at] = (static_cast<uint64_t>(tileBase + ty * tilesX + tx) << 32) | depthBits;
                if constexpr (EMIT_INDEX) {
                    values[at] = at;
                    pairSlot[at] = s;
                } else {
                    values[at] = s;
                }
            } else {
                atomicAdd(overflow, 1u);
            }
            written += 1;
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
    duplicateSlot<EMIT_INDEX>(s, f.width, f.height, f.tilesX, 0u, proj, rects, touched, offsets, keys, values,
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
    const ViewInfo& v = views[s / splatsPerView];
    duplicateSlot<false>(s, v.width, v.height, v.tilesX, v.tileBase, proj, rects, touched, offsets, keys, values,
                         nullptr, capacity, overflow);
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
