// Exclusive u32 prefix sum in one pass: each workgroup scans a 4096-element
// partition in registers and LDS, publishes its aggregate, and looks back over
// its predecessors' published aggregates/prefixes (decoupled look-back) for its
// starting offset. Partitions are numbered by workgroup id: workgroups reach
// each XCD in ascending id order, so the lowest unfinished partition is always
// resident and every look-back ends (an atomic ticket would cost a
// same-address device atomic, tens of nanoseconds across XCDs, a partition).
//
// Replaces the reference's three-dispatch chunked scan (modules/gpu/src/algo/
// PrefixSum.cpp:58-88, shaders/lrt/algo/prefix_*.slang), whose chunk starts
// are computed by a single thread. Same contract: out[i] = in[0] + ... +
// in[i-1] in u32 arithmetic, total = the sum of all.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "kernels.h"

namespace lrt::kern {
namespace {

constexpr uint32_t kBlock = 256;
constexpr uint32_t kItems = 16;
constexpr uint32_t kPart = kBlock * kItems;
constexpr uint64_t kAggregate = 1ull << 62;
constexpr uint64_t kInclusive = 2ull << 62;
constexpr uint64_t kValueMask = (1ull << 62) - 1;
constexpr uint32_t kSpinLimit = 1u << 18;

// Zeroes words [0, n0) of a and [0, n1) of b: two small clears in one launch.
__global__ void zero2Kernel(uint32_t* a, uint64_t n0, uint32_t* b, uint64_t n1) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n0 + n1; i += stride) {
        if (i < n0) {
            a[i] = 0u;
        } else {
            b[i - n0] = 0u;
        }
    }
}

__global__ void fill32Kernel(uint32_t* dst, uint32_t value, uint64_t count) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += stride) {
        dst[i] = value;
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

__global__ __launch_bounds__(kBlock) void scanKernel(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                     uint32_t n, uint64_t* __restrict__ status,
                                                     uint32_t* __restrict__ total,
                                                     uint32_t* __restrict__ nonzeroParts, uint32_t* __restrict__ error,
                                                     bool vectorised) {
    __shared__ uint32_t sWave[kBlock / 64];
    __shared__ uint32_t sNonzero[kBlock / 64];
    __shared__ uint32_t sPrefix;
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u;
    const uint32_t wave = tid >> 6;
    const uint32_t part = blockIdx.x;
    const uint64_t first = static_cast<uint64_t>(part) * kPart + static_cast<uint64_t>(tid) * kItems;

    uint32_t v[kItems];
    if (vectorised && first + kItems <= n) {
        const uint4* src = reinterpret_cast<const uint4*>(in + first);
#pragma unroll
        for (uint32_t q = 0; q < kItems / 4; ++q) {
            const uint4 w = src[q];
            v[q * 4 + 0] = w.x;
            v[q * 4 + 1] = w.y;
            v[q * 4 + 2] = w.z;
            v[q * 4 + 3] = w.w;
        }
    } else {
#pragma unroll
        for (uint32_t j = 0; j < kItems; ++j) v[j] = first + j < n ? in[first + j] : 0u;
    }
    uint32_t sum = 0, nz = 0;
#pragma unroll
    for (uint32_t j = 0; j < kItems; ++j) {
        sum += v[j];
        nz += v[j] != 0u ? 1u : 0u;
    }
    const uint32_t inclusive = waveInclusiveScan(sum, lane);
    const uint32_t nzWave = waveInclusiveScan(nz, lane);
    if (lane == 63) {
        sWave[wave] = inclusive;
        sNonzero[wave] = nzWave;
    }
    __syncthreads();
    uint32_t waveBase = 0, blockTotal = 0;
#pragma unroll
    for (uint32_t w = 0; w < kBlock / 64; ++w) {
        const uint32_t t = sWave[w];
        if (w < wave) waveBase += t;
        blockTotal += t;
    }
    if (tid == 0) nonzeroParts[part] = sNonzero[0] + sNonzero[1] + sNonzero[2] + sNonzero[3];

    // Decoupled look-back by the first wave: publish this partition's
    // aggregate, then read 64 predecessors' words at once (lane 0 the nearest;
    // before partition 0 reads as an inclusive 0), wait until every word up to
    // the nearest inclusive one is published, and add them; with no inclusive
    // word among the 64, move the window back by 64.
    if (wave == 0) {
        uint64_t* mine = status + part;
        uint64_t prefix = 0;   // 64-bit, so that a sum past 2^32 shows as such
        if (part == 0) {
            if (lane == 0) __hip_atomic_store(mine, kInclusive | blockTotal, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            if (lane == 0) __hip_atomic_store(mine, kAggregate | blockTotal, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            int64_t end = static_cast<int64_t>(part);   // the window is [end - 64, end)
            uint32_t spins = 0;
            while (true) {
                const int64_t p = end - 1 - static_cast<int64_t>(lane);
                const uint64_t s = p >= 0 ? __hip_atomic_load(status + p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                          : kInclusive;
                const uint64_t flag = s & ~kValueMask;
                const uint64_t inclusive = __ballot(flag == kInclusive);
                const uint64_t pending = __ballot(flag == 0);
                const uint32_t nearest = inclusive != 0 ? static_cast<uint32_t>(__builtin_ctzll(inclusive)) : 64u;
                const uint64_t needed = nearest >= 63 ? ~0ull : ((2ull << nearest) - 1ull);
                if ((pending & needed) != 0) {
                    // A predecessor (a lower workgroup id) is running or next on
                    // its XCD. The bound only guards a broken launch.
                    if (++spins > kSpinLimit) {
                        if (lane == 0) atomicOr(error, 1u);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                uint64_t take = lane <= nearest ? (s & kValueMask) : 0ull;
#pragma unroll
                for (uint32_t d = 1; d < 64; d <<= 1) take += __shfl_xor(take, static_cast<int>(d), 64);
                prefix += take;
                if (nearest < 64) break;
                end -= 64;
            }
            if (lane == 0) {
                __hip_atomic_store(mine, kInclusive | ((prefix + blockTotal) & kValueMask), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        if (lane == 0) {
            sPrefix = static_cast<uint32_t>(prefix);
            // The total saturates at 2^32 - 1 rather than wrap.
            const uint64_t all = prefix + blockTotal;
            if (static_cast<uint64_t>(part + 1) * kPart >= n) *total = all > 0xFFFFFFFFull ? 0xFFFFFFFFu : static_cast<uint32_t>(all);
        }
    }
    __syncthreads();
    uint32_t run = sPrefix + waveBase + inclusive - sum;
    if (vectorised && first + kItems <= n) {
        uint4* dst = reinterpret_cast<uint4*>(out + first);
#pragma unroll
        for (uint32_t q = 0; q < kItems / 4; ++q) {
            uint4 w;
            w.x = run; run += v[q * 4 + 0];
            w.y = run; run += v[q * 4 + 1];
            w.z = run; run += v[q * 4 + 2];
            w.w = run; run += v[q * 4 + 3];
            dst[q] = w;
        }
    } else {
#pragma unroll
        for (uint32_t j = 0; j < kItems; ++j) {
            if (first + j < n) out[first + j] = run;
            run += v[j];
        }
    }
}

// One workgroup: the partitions' non-zero counts added up; with depthRange
// non-null, the depth-key entries (two words each: least, greatest) folded.
__global__ __launch_bounds__(1024) void sumPartsKernel(const uint32_t* __restrict__ parts, uint32_t count,
                                                       uint32_t* __restrict__ out,
                                                       const uint2* __restrict__ depthBlocks, uint32_t depthEntries,
                                                       uint32_t* __restrict__ depthRange) {
    __shared__ uint32_t sWave[3][1024 / 64];
    uint32_t sum = 0;
    for (uint32_t i = threadIdx.x; i < count; i += 1024) sum += parts[i];
    uint32_t lo = 0xFFFFFFFFu, hi = 0u;
    if (depthRange != nullptr) {
        constexpr uint32_t kBatch = 8;   // loads in flight before the first fold
        for (uint32_t e0 = threadIdx.x; e0 < depthEntries; e0 += 1024 * kBatch) {
            uint2 r[kBatch];
#pragma unroll
            for (uint32_t b = 0; b < kBatch; ++b) {
                const uint32_t e = e0 + b * 1024;
                r[b] = e < depthEntries ? depthBlocks[e] : make_uint2(0xFFFFFFFFu, 0u);
            }
#pragma unroll
            for (uint32_t b = 0; b < kBatch; ++b) {
                lo = min(lo, r[b].x);
                hi = max(hi, r[b].y);
            }
        }
    }
#pragma unroll
    for (uint32_t d = 1; d < 64; d <<= 1) {
        sum += static_cast<uint32_t>(__shfl_xor(static_cast<int>(sum), static_cast<int>(d), 64));
        lo = min(lo, static_cast<uint32_t>(__shfl_xor(static_cast<int>(lo), static_cast<int>(d), 64)));
        hi = max(hi, static_cast<uint32_t>(__shfl_xor(static_cast<int>(hi), static_cast<int>(d), 64)));
    }
    if ((threadIdx.x & 63u) == 0) {
        sWave[0][threadIdx.x >> 6] = sum;
        sWave[1][threadIdx.x >> 6] = lo;
        sWave[2][threadIdx.x >> 6] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t all = 0, allLo = 0xFFFFFFFFu, allHi = 0u;
        for (uint32_t w = 0; w < 1024 / 64; ++w) {
            all += sWave[0][w];
            allLo = min(allLo, sWave[1][w]);
            allHi = max(allHi, sWave[2][w]);
        }
        *out = all;
        if (depthRange != nullptr) {
            depthRange[0] = allLo;
            depthRange[1] = allHi;
        }
    }
}

}   // namespace

void fill32(uint32_t* dst, uint32_t value, uint64_t count, hipStream_t stream) {
    if (count == 0) return;
    const uint64_t blocks = (count + 255) / 256;
    const uint32_t grid = static_cast<uint32_t>(blocks < 4096 ? blocks : 4096);
    fill32Kernel<<<grid, 256, 0, stream>>>(dst, value, count);
}

uint32_t scanPartitions(uint32_t n) { return (n + kPart - 1) / kPart; }

// ---- the prefix sum that allocates nothing ------------------------------------
// Its look-back words, per-partition counts and counters live in one array in
// the code object, each device holding its own copy: the status words of
// kStaticScanMax elements, then one word holding the non-zero count (low u32)
// and the error flags (high u32).
namespace {
constexpr uint64_t kStaticStatusWords = uint64_t{(kStaticScanMax + kPart - 1) / kPart} * 3 / 2 + 2;
constexpr uint64_t kStaticScratchWords = kStaticStatusWords + 1;
static __device__ __align__(256) uint64_t scratch[kStaticScratchWords];

constexpr int kStaticDevices = 64;
std::mutex gScratchMutex;
uint64_t* gScratch[kStaticDevices] = {};   // each device's address of `scratch`

// The current device's address of `scratch`, looked up once per device with
// that device current.
hipError_t staticScratch(uint64_t** out) {
    int device = 0;
    hipError_t e = hipGetDevice(&device);
    if (e != hipSuccess) return e;
    if (device < 0 || device >= kStaticDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(gScratchMutex);
    if (gScratch[device] == nullptr) {
        void* ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        gScratch[device] = static_cast<uint64_t*>(ptr);
    }
    *out = gScratch[device];
    return hipSuccess;
}
}   // namespace

hipError_t scanExclusiveStatic(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* total, hipStream_t stream) {
    if (n > kStaticScanMax) return hipErrorInvalidValue;
    uint64_t* ptr = nullptr;
    const hipError_t e = staticScratch(&ptr);
    if (e != hipSuccess) return e;
    uint32_t* counters = reinterpret_cast<uint32_t*>(ptr + kStaticStatusWords);
    // The error flags are OR-ed by the scan: cleared with the look-back words.
    scanExclusive(in, out, n, ptr, total, counters, counters + 1, stream, nullptr, 0, nullptr, counters + 1, 1);
    return hipGetLastError();
}

void scanExclusive(const uint32_t* in, uint32_t* out, uint32_t n, uint64_t* status, uint32_t* total,
                   uint32_t* nonzero, uint32_t* error, hipStream_t stream, const uint32_t* depthBlocks,
                   uint32_t depthEntries, uint32_t* depthRange, uint32_t* alsoZero, uint32_t alsoZeroWords,
                   bool zeroed) {
    const uint32_t parts = scanPartitions(n);
    if (const uint64_t words = static_cast<uint64_t>(parts) * 2 + alsoZeroWords; words > 0 && !zeroed) {
        // (an empty scan with nothing else to clear launches nothing: a grid of 0 is an error)
        const uint32_t grid = static_cast<uint32_t>(std::min<uint64_t>((words + 255) / 256, 4096));
        zero2Kernel<<<grid, 256, 0, stream>>>(reinterpret_cast<uint32_t*>(status), static_cast<uint64_t>(parts) * 2,
                                              alsoZero, alsoZero != nullptr ? alsoZeroWords : 0u);
    }
    if (n == 0) {
        fill32(total, 0u, 1, stream);
        fill32(nonzero, 0u, 1, stream);
        if (depthRange != nullptr) {
            fill32(depthRange, 0xFFFFFFFFu, 1, stream);
            fill32(depthRange + 1, 0u, 1, stream);
        }
        return;
    }
    // Each partition's non-zero count after its look-back words.
    uint32_t* nonzeroParts = reinterpret_cast<uint32_t*>(status + parts);
    // 16-byte loads and stores only where both arrays allow them.
    const bool vectorised = (reinterpret_cast<uintptr_t>(in) % 16 == 0) && (reinterpret_cast<uintptr_t>(out) % 16 == 0);
    scanKernel<<<parts, kBlock, 0, stream>>>(in, out, n, status, total, nonzeroParts, error, vectorised);
    sumPartsKernel<<<1, 1024, 0, stream>>>(nonzeroParts, parts, nonzero,
                                           reinterpret_cast<const uint2*>(depthBlocks), depthEntries, depthRange);
}

}   // namespace lrt::kern
