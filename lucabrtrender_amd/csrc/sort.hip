// Stable LSD radix sort of (u64 key, u32 value) pairs: one histogram pass over
// the keys for every digit at once, then one onesweep launch per 8-bit digit.
//
// A onesweep workgroup (256 threads, 4 wave64s) takes the next partition of
// 3840 pairs by atomic ticket and
//   1. ranks its keys by digit within each wave with ballot-built match masks
//      (eight __ballot per 64 keys), keeping a per-wave digit histogram in LDS;
//   2. turns the four wave histograms into per-wave offsets and a block digit
//      count, publishes the count for its partition and looks back over the
//      earlier partitions' published counts/prefixes (decoupled look-back) for
//      the number of equal digits before it;
//   3. scatters its pairs into LDS in local digit order and writes each digit's
//      run to global memory contiguously.
// Ranking in (wave, key slot, lane) order, which is input order, and
// partitions taken in input order make every pass stable, so the sort is.
//
// Replaces the reference's chunk-serial LSD radix (modules/gpu/src/algo/
// RadixSort.cpp:80-158, shaders/lrt/algo/radix_*.slang) with the same result:
// a stable sort on the chosen key bits.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kernels.h"

namespace lrt::kern {
namespace {

constexpr uint32_t kBlock = 256;
constexpr uint32_t kWaves = kBlock / 64;
constexpr uint32_t kKpt = 15;
constexpr uint32_t kPart = kBlock * kKpt;   // 3840 pairs a partition
constexpr uint32_t kRadix = 256;
constexpr uint32_t kMaxPasses = 8;
constexpr uint32_t kAggregate = 1u << 30;
constexpr uint32_t kInclusive = 2u << 30;
constexpr uint32_t kValueMask = (1u << 30) - 1;
constexpr uint32_t kSpinLimit = 1u << 22;

// The pairs to sort: the host's count, or the one a kernel left in device
// memory (then the host's is the capacity and bounds it).
__device__ __forceinline__ uint32_t pairCount(uint32_t n, const uint32_t* __restrict__ nDevice) {
    if (nDevice == nullptr) return n;
    const uint32_t d = *nDevice;
    return d < n ? d : n;
}

template <KeyMode MODE>
__device__ __forceinline__ uint64_t sortKey(uint64_t key, uint32_t bitLo, uint64_t mask) {
    if constexpr (MODE == KeyMode::TileDepth) {
        return (((key >> 32) << 24) | ((static_cast<uint32_t>(key) & 0x7FFFFFFFu) >> 7)) & mask;
    } else {
        return (key >> bitLo) & mask;
    }
}

template <KeyMode MODE>
__global__ __launch_bounds__(kBlock) void histogramKernel(const uint64_t* __restrict__ keys, uint32_t cap,
                                                          const uint32_t* __restrict__ nDevice, uint32_t bitLo,
                                                          uint64_t mask, uint32_t passes,
                                                          uint32_t* __restrict__ hist) {
    __shared__ uint32_t sHist[kMaxPasses * kRadix];
    const uint32_t n = pairCount(cap, nDevice);
    for (uint32_t j = threadIdx.x; j < passes * kRadix; j += kBlock) sHist[j] = 0;
    __syncthreads();
    const uint32_t stride = gridDim.x * kBlock;
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
        const uint64_t k = sortKey<MODE>(keys[i], bitLo, mask);
        for (uint32_t p = 0; p < passes; ++p) {
            atomicAdd(&sHist[p * kRadix + static_cast<uint32_t>((k >> (8 * p)) & 255u)], 1u);
        }
    }
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < passes * kRadix; j += kBlock) {
        const uint32_t c = sHist[j];
        if (c != 0) atomicAdd(&hist[j], c);
    }
}

// Clears the look-back words of the partitions the count needs, every pass.
__global__ __launch_bounds__(kBlock) void clearStatusKernel(uint32_t* __restrict__ status, uint32_t passes,
                                                            uint32_t capParts, uint32_t cap,
                                                            const uint32_t* __restrict__ nDevice) {
    const uint32_t n = pairCount(cap, nDevice);
    const uint64_t used = static_cast<uint64_t>((n + kPart - 1) / kPart) * kRadix;
    const uint64_t total = used * passes;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; i < total; i += stride) {
        const uint64_t pass = i / used, at = i - pass * used;
        status[pass * capParts * kRadix + at] = 0u;
    }
}

__device__ __forceinline__ uint32_t waveInclusiveScan(uint32_t x, uint32_t lane) {
#pragma unroll
    for (uint32_t d = 1; d < 64; d <<= 1) {
        const uint32_t t = __shfl_up(x, d, 64);
        if (lane >= d) x += t;
    }
    return x;
}

// One workgroup per pass: exclusive scan of the 256 digit counts.
__global__ __launch_bounds__(kRadix) void digitOffsetsKernel(const uint32_t* __restrict__ hist,
                                                             uint32_t* __restrict__ offsets) {
    __shared__ uint32_t sWave[kRadix / 64];
    const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
    const uint32_t c = hist[blockIdx.x * kRadix + t];
    const uint32_t incl = waveInclusiveScan(c, lane);
    if (lane == 63) sWave[wave] = incl;
    __syncthreads();
    uint32_t base = 0;
    for (uint32_t w = 0; w < wave; ++w) base += sWave[w];
    offsets[blockIdx.x * kRadix + t] = base + incl - c;
}

template <KeyMode MODE>
__global__ __launch_bounds__(kBlock) void onesweepKernel(const uint64_t* __restrict__ keysIn,
                                                         const uint32_t* __restrict__ valsIn,
                                                         uint64_t* __restrict__ keysOut,
                                                         uint32_t* __restrict__ valsOut, uint32_t cap,
                                                         const uint32_t* __restrict__ nDevice,
                                                         uint32_t bitLo, uint64_t mask, uint32_t shift,
                                                         const uint32_t* __restrict__ digitOffsets,
                                                         uint32_t* __restrict__ status,
                                                         uint32_t* __restrict__ ticket,
                                                         uint32_t* __restrict__ error) {
    __shared__ uint64_t sKeys[kPart];
    __shared__ uint32_t sVals[kPart];
    __shared__ uint32_t sWave[kWaves][kRadix];
    __shared__ uint32_t sLocal[kRadix];
    __shared__ uint32_t sGlobal[kRadix];
    __shared__ uint32_t sScan[kWaves];
    __shared__ uint32_t sPart;

    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u;
    const uint32_t wave = tid >> 6;
    if (tid == 0) sPart = atomicAdd(ticket, 1u);
#pragma unroll
    for (uint32_t w = 0; w < kWaves; ++w) sWave[w][tid] = 0;
    __syncthreads();
    const uint32_t n = pairCount(cap, nDevice);
    const uint32_t part = sPart;
    const uint32_t base = part * kPart;
    // Launched for the capacity: partitions past the count have nothing to
    // do, and no partition before them waits on them.
    if (base >= n) return;
    const uint32_t waveFirst = base + wave * (kKpt * 64);

    uint64_t k[kKpt];
    uint32_t v[kKpt];
    uint32_t d[kKpt];
    uint32_t r[kKpt];
#pragma unroll
    for (uint32_t j = 0; j < kKpt; ++j) {
        const uint32_t idx = waveFirst + j * 64 + lane;
        const bool valid = idx < n;
        k[j] = valid ? keysIn[idx] : 0ull;
        v[j] = valid ? valsIn[idx] : 0u;
        d[j] = static_cast<uint32_t>((sortKey<MODE>(k[j], bitLo, mask) >> shift) & 255u);
    }

    // Rank within the wave: peers = lanes holding the same digit.
    const uint64_t below = (1ull << lane) - 1ull;
#pragma unroll
    for (uint32_t j = 0; j < kKpt; ++j) {
        const bool valid = waveFirst + j * 64 + lane < n;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (uint32_t b = 0; b < 8; ++b) {
            const bool bit = ((d[j] >> b) & 1u) != 0u;
            const uint64_t vote = __ballot(bit);
            peers &= bit ? vote : ~vote;
        }
        const uint32_t before = static_cast<uint32_t>(__popcll(peers & below));
        const uint32_t count = static_cast<uint32_t>(__popcll(peers));
        // The count a leader stores is read by other lanes of the wave at the
        // next key: relaxed atomics, so the compiler keeps no per-lane copy.
        uint32_t prev = 0;
        if (valid) prev = __hip_atomic_load(&sWave[wave][d[j]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        r[j] = prev + before;
        if (valid && before == 0) {
            __hip_atomic_store(&sWave[wave][d[j]], prev + count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
    }
    __syncthreads();

    // Digit tid: wave offsets, block count, publication and look-back.
    const uint32_t c0 = sWave[0][tid], c1 = sWave[1][tid], c2 = sWave[2][tid], c3 = sWave[3][tid];
    sWave[0][tid] = 0;
    sWave[1][tid] = c0;
    sWave[2][tid] = c0 + c1;
    sWave[3][tid] = c0 + c1 + c2;
    const uint32_t total = c0 + c1 + c2 + c3;
    uint32_t* mine = status + static_cast<uint64_t>(part) * kRadix + tid;
    __hip_atomic_store(mine, (part == 0 ? kInclusive : kAggregate) | total, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t incl = waveInclusiveScan(total, lane);
    if (lane == 63) sScan[wave] = incl;
    uint32_t prefix = 0;
    if (part > 0) {
        int64_t p = static_cast<int64_t>(part) - 1;
        uint32_t spins = 0;
        while (p >= 0) {
            const uint32_t s = __hip_atomic_load(status + static_cast<uint64_t>(p) * kRadix + tid, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT);
            const uint32_t flag = s & ~kValueMask;
            if (flag == 0) {
                // The partition waited on took its ticket earlier and is
                // running; the bound only guards a broken launch.
                if (++spins > kSpinLimit) {
                    atomicOr(error, 1u);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
                continue;
            }
            prefix += s & kValueMask;
            if (flag == kInclusive) break;
            --p;
        }
        __hip_atomic_store(mine, kInclusive | (prefix + total), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    uint32_t waveBase = 0;
    for (uint32_t w = 0; w < wave; ++w) waveBase += sScan[w];
    const uint32_t localStart = waveBase + incl - total;
    sLocal[tid] = localStart;
    sGlobal[tid] = digitOffsets[tid] + prefix - localStart;
    __syncthreads();

    // Local scatter into digit order, then contiguous runs out.
#pragma unroll
    for (uint32_t j = 0; j < kKpt; ++j) {
        if (waveFirst + j * 64 + lane < n) {
            const uint32_t at = sLocal[d[j]] + sWave[wave][d[j]] + r[j];
            if (at < kPart) {
                sKeys[at] = k[j];
                sVals[at] = v[j];
            } else {
                atomicOr(error, 2u);
            }
        }
    }
    __syncthreads();
    const uint32_t count = n - base < kPart ? n - base : kPart;
    for (uint32_t i = tid; i < count; i += kBlock) {
        const uint64_t key = sKeys[i];
        const uint32_t dd = static_cast<uint32_t>((sortKey<MODE>(key, bitLo, mask) >> shift) & 255u);
        const uint32_t out = sGlobal[dd] + i;
        if (out < n) {
            keysOut[out] = key;
            valsOut[out] = sVals[i];
        } else {
            atomicOr(error, 4u);
        }
    }
}

template <KeyMode MODE>
void launchPasses(uint64_t* keys, uint32_t* values, uint64_t* keysAlt, uint32_t* valuesAlt, uint32_t n,
                  const uint32_t* nDevice, uint32_t bitLo, uint32_t keyBits, const SortWorkspace& ws,
                  hipStream_t stream, hipEvent_t afterHistogram) {
    const uint32_t passes = sortPasses(keyBits);
    const uint32_t parts = sortPartitions(n);
    const uint64_t mask = keyBits >= 64 ? ~0ull : ((1ull << keyBits) - 1ull);
    fill32(ws.histogram, 0u, static_cast<uint64_t>(passes) * kRadix, stream);
    if (nDevice == nullptr) {
        fill32(ws.status, 0u, static_cast<uint64_t>(passes) * parts * kRadix, stream);
    } else {
        const uint64_t words = static_cast<uint64_t>(passes) * parts * kRadix;
        const uint32_t grid = static_cast<uint32_t>(std::min<uint64_t>((words + kBlock - 1) / kBlock, 4096));
        clearStatusKernel<<<grid, kBlock, 0, stream>>>(ws.status, passes, parts, n, nDevice);
    }
    fill32(ws.tickets, 0u, passes, stream);
    const uint32_t histBlocks = parts < 1024 ? parts : 1024;
    histogramKernel<MODE><<<histBlocks, kBlock, 0, stream>>>(keys, n, nDevice, bitLo, mask, passes, ws.histogram);
    digitOffsetsKernel<<<passes, kRadix, 0, stream>>>(ws.histogram, ws.offsets);
    if (afterHistogram != nullptr) (void)hipEventRecord(afterHistogram, stream);
    const uint64_t* kin = keys;
    const uint32_t* vin = values;
    uint64_t* kout = keysAlt;
    uint32_t* vout = valuesAlt;
    for (uint32_t p = 0; p < passes; ++p) {
        onesweepKernel<MODE><<<parts, kBlock, 0, stream>>>(
            kin, vin, kout, vout, n, nDevice, bitLo, mask, 8 * p, ws.offsets + p * kRadix,
            ws.status + static_cast<uint64_t>(p) * parts * kRadix, ws.tickets + p, ws.error);
        const uint64_t* nk = kout;
        const uint32_t* nv = vout;
        kout = const_cast<uint64_t*>(kin);
        vout = const_cast<uint32_t*>(vin);
        kin = nk;
        vin = nv;
    }
}

}   // namespace

uint32_t sortPartitions(uint32_t n) { return (n + kPart - 1) / kPart; }
uint32_t sortPasses(uint32_t keyBits) { return (keyBits + 7) / 8; }
uint64_t sortWorkspaceWords(uint32_t n, uint32_t keyBits) {
    const uint64_t passes = sortPasses(keyBits);
    return passes * kRadix * 2 + passes * sortPartitions(n) * kRadix + passes + 1;
}

bool sortPairs(uint64_t* keys, uint32_t* values, uint64_t* keysAlt, uint32_t* valuesAlt, uint32_t n, KeyMode mode,
               uint32_t bitLo, uint32_t keyBits, const SortWorkspace& ws, hipStream_t stream,
               hipEvent_t afterHistogram, const uint32_t* nDevice) {
    if ((nDevice == nullptr && n < 2) || n == 0 || keyBits == 0) {
        if (afterHistogram != nullptr) (void)hipEventRecord(afterHistogram, stream);
        return false;
    }
    if (mode == KeyMode::TileDepth) {
        launchPasses<KeyMode::TileDepth>(keys, values, keysAlt, valuesAlt, n, nDevice, bitLo, keyBits, ws, stream,
                                         afterHistogram);
    } else {
        launchPasses<KeyMode::Plain>(keys, values, keysAlt, valuesAlt, n, nDevice, bitLo, keyBits, ws, stream,
                                     afterHistogram);
    }
    return (sortPasses(keyBits) & 1u) != 0u;
}

}   // namespace lrt::kern
