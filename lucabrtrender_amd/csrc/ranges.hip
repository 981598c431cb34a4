// Tile ranges: where each tile's run of sorted pairs starts and ends. Every
// boundary is written by exactly one thread; tiles with no pairs keep (0, 0).
// The reference's splatTilesClear + splatRanges (shaders/lrt/splat/
// splat_tiles_clear.slang:8-17, splat_ranges.slang:11-25), with the tile read
// from the high word of the 64-bit key.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace lrt::kern {
namespace {

constexpr uint32_t kBlock = 256;

__global__ __launch_bounds__(kBlock) void rangesKernel(const uint64_t* __restrict__ keys, uint32_t cap,
                                                       const uint32_t* __restrict__ nDevice, uint32_t tiles,
                                                       uint2* __restrict__ ranges) {
    uint32_t pairs = cap;
    if (nDevice != nullptr && *nDevice < cap) pairs = *nDevice;
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= pairs) return;
    const uint32_t t = static_cast<uint32_t>(keys[i] >> 32);
    if (t >= tiles) return;
    if (i == 0 || static_cast<uint32_t>(keys[i - 1] >> 32) != t) ranges[t].x = i;
    if (i + 1 == pairs || static_cast<uint32_t>(keys[i + 1] >> 32) != t) ranges[t].y = i + 1;
}

}   // namespace

void tileRanges(const uint64_t* keys, uint32_t pairs, uint32_t tiles, uint2* ranges, hipStream_t stream,
                const uint32_t* nDevice) {
    fill32(reinterpret_cast<uint32_t*>(ranges), 0u, static_cast<uint64_t>(tiles) * 2, stream);
    if (pairs == 0) return;
    rangesKernel<<<(pairs + kBlock - 1) / kBlock, kBlock, 0, stream>>>(keys, pairs, nDevice, tiles, ranges);
}

}   // namespace lrt::kern
