// Tile ranges: where each tile's run of sorted pairs starts and ends. Every
// boundary is written by exactly one thread; tiles with no pairs keep (0, 0).
// The reference's splatTilesClear + splatRanges (shaders/lrt/splat/
// splat_tiles_clear.slang:8-17, splat_ranges.slang:11-25), with the tile read
// from the high word of the 64-bit key or from a narrowed u32 key.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace lrt::kern {
namespace {

constexpr uint32_t kBlock = 256;
constexpr uint32_t kPer = 4;                     // consecutive pairs a thread
constexpr uint32_t kNone = 0xFFFFFFFFu;          // no tile: before the first pair, after the last

template <int DPP_CTRL>
__device__ __forceinline__ uint32_t fromLane(uint32_t v, uint32_t missing) {
    return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(static_cast<int>(missing), static_cast<int>(v), DPP_CTRL,
                                                             0xF, 0xF, false));
}

// The tile of a sorted key: the high word of a u64 key, or a narrowed u32
// key shifted right by tileShift.
__device__ __forceinline__ uint32_t tileOf(uint64_t key, uint32_t) { return static_cast<uint32_t>(key >> 32); }
__device__ __forceinline__ uint32_t tileOf(uint32_t key, uint32_t tileShift) { return key >> tileShift; }

// Thread g takes pairs [4g, 4g + 4) with 16-byte loads (two of u64 keys, one
// of u32); the tile before its first pair and after its last come from the
// neighbouring lanes (DPP wave shifts), the wave's two end lanes loading
// theirs. Keys must be 16-byte aligned (the rasteriser's sort buffers are).
template <typename Key>
__global__ __launch_bounds__(kBlock) void rangesKernel(const Key* __restrict__ keys, uint32_t tileShift, uint32_t cap,
                                                       const uint32_t* __restrict__ nDevice, uint32_t tiles,
                                                       uint2* __restrict__ ranges) {
    uint32_t pairs = cap;
    if (nDevice != nullptr && *nDevice < cap) pairs = *nDevice;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t i0 = (blockIdx.x * kBlock + threadIdx.x) * kPer;
    // Launched for the capacity: a wave wholly past the count leaves at once.
    if (i0 - lane * kPer >= pairs) return;
    uint32_t t[kPer];
    if (i0 + kPer <= pairs) {
        if constexpr (sizeof(Key) == 8) {
            const uint4* k4 = reinterpret_cast<const uint4*>(keys + i0);
            const uint4 a = k4[0], b = k4[1];
            t[0] = a.y;
            t[1] = a.w;
            t[2] = b.y;
            t[3] = b.w;
        } else {
            const uint4 a = *reinterpret_cast<const uint4*>(keys + i0);
            t[0] = a.x >> tileShift;
            t[1] = a.y >> tileShift;
            t[2] = a.z >> tileShift;
            t[3] = a.w >> tileShift;
        }
    } else {
#pragma unroll
        for (uint32_t j = 0; j < kPer; ++j) t[j] = i0 + j < pairs ? tileOf(keys[i0 + j], tileShift) : kNone;
    }
    uint32_t before = fromLane<0x138>(t[kPer - 1], kNone);   // wave_shr:1, lane - 1's last
    uint32_t after = fromLane<0x130>(t[0], kNone);           // wave_shl:1, lane + 1's first
    if (lane == 0) before = i0 > 0 && i0 - 1 < pairs ? tileOf(keys[i0 - 1], tileShift) : kNone;
    if (lane == 63) after = i0 + kPer < pairs ? tileOf(keys[i0 + kPer], tileShift) : kNone;
#pragma unroll
    for (uint32_t j = 0; j < kPer; ++j) {
        const uint32_t tile = t[j];
        const uint32_t prev = j == 0 ? before : t[j - 1];
        const uint32_t next = j + 1 == kPer ? after : t[j + 1];
        if (tile >= tiles) continue;   // also kNone
        if (prev != tile) ranges[tile].x = i0 + j;
        if (next != tile) ranges[tile].y = i0 + j + 1;
    }
}

__global__ __launch_bounds__(kBlock) void widenKeysKernel(const uint32_t* __restrict__ keys, uint32_t tileShift,
                                                          const uint32_t* __restrict__ slots,
                                                          const float4* __restrict__ proj, uint32_t pairs,
                                                          uint64_t* __restrict__ out) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= pairs) return;
    const uint64_t tile = keys[i] >> tileShift;
    const float z = proj[static_cast<uint64_t>(slots[i]) * 3].z;
    out[i] = tile << 32 | __float_as_uint(z);
}

}   // namespace

namespace {
template <typename Key>
void launchRanges(const Key* keys, uint32_t tileShift, uint32_t pairs, uint32_t tiles, uint2* ranges,
                  hipStream_t stream, const uint32_t* nDevice, bool zeroed) {
    if (!zeroed) fill32(reinterpret_cast<uint32_t*>(ranges), 0u, static_cast<uint64_t>(tiles) * 2, stream);
    if (pairs == 0) return;
    const uint64_t threads = (uint64_t{pairs} + kPer - 1) / kPer;
    rangesKernel<Key><<<static_cast<uint32_t>((threads + kBlock - 1) / kBlock), kBlock, 0, stream>>>(
        keys, tileShift, pairs, nDevice, tiles, ranges);
}
}   // namespace

void tileRanges(const uint64_t* keys, uint32_t pairs, uint32_t tiles, uint2* ranges, hipStream_t stream,
                const uint32_t* nDevice, bool zeroed) {
    launchRanges(keys, 32u, pairs, tiles, ranges, stream, nDevice, zeroed);
}

void tileRangesNarrow(const uint32_t* keys, uint32_t tileShift, uint32_t pairs, uint32_t tiles, uint2* ranges,
                      hipStream_t stream, const uint32_t* nDevice, bool zeroed) {
    launchRanges(keys, tileShift, pairs, tiles, ranges, stream, nDevice, zeroed);
}

void widenKeys(const uint32_t* keys, uint32_t tileShift, const uint32_t* slots, const float4* proj, uint32_t pairs,
               uint64_t* out, hipStream_t stream) {
    if (pairs == 0) return;
    widenKeysKernel<<<(pairs + kBlock - 1) / kBlock, kBlock, 0, stream>>>(keys, tileShift, slots, proj, pairs, out);
}

}   // namespace lrt::kern
