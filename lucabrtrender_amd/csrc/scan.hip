// Exclusive u32 prefix sum in one pass: each workgroup scans a 4096-element
// partition in registers and LDS, publishes its aggregate, and looks back over
// its predecessors' published aggregates/prefixes (decoupled look-back) for its
// starting offset. Partitions are numbered by an atomic ticket, so a
// workgroup only ever waits on partitions that were already running.
//
// Replaces the reference's three-dispatch chunked scan (modules/gpu/src/algo/
// PrefixSum.cpp:58-88, shaders/lrt/algo/prefix_*.slang), whose chunk starts
// are computed by a single thread. Same contract: out[i] = in[0] + ... +
// in[i-1] in u32 arithmetic, total = the sum of all.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace lrt::kern {
namespace {

constexpr uint32_t kBlock = 256;
constexpr uint32_t kItems = 16;
constexpr uint32_t kPart = kBlock * kItems;
constexpr uint64_t kAggregate = 1ull << 62;
constexpr uint64_t kInclusive = 2ull << 62;
constexpr uint64_t kValueMask = (1ull << 62) - 1;
constexpr uint32_t kSpinLimit = 1u << 22;

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
                                                     uint32_t* __restrict__ ticket, uint32_t* __restrict__ total,
                                                     uint32_t* __restrict__ nonzero, uint32_t* __restrict__ error,
                                                     bool vectorised) {
    __shared__ uint32_t sPart;
    __shared__ uint32_t sWave[kBlock / 64];
    __shared__ uint32_t sPrefix;
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u;
    const uint32_t wave = tid >> 6;
    if (tid == 0) sPart = atomicAdd(ticket, 1u);
    __syncthreads();
    const uint32_t part = sPart;
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
    if (lane == 63) sWave[wave] = inclusive;
    __syncthreads();
    uint32_t waveBase = 0, blockTotal = 0;
#pragma unroll
    for (uint32_t w = 0; w < kBlock / 64; ++w) {
        const uint32_t t = sWave[w];
        if (w < wave) waveBase += t;
        blockTotal += t;
    }
    if (lane == 63) atomicAdd(nonzero, nzWave);

    if (tid == 0) {
        uint64_t* mine = status + part;
        uint32_t prefix = 0;
        if (part == 0) {
            __hip_atomic_store(mine, kInclusive | blockTotal, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            __hip_atomic_store(mine, kAggregate | blockTotal, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            int64_t p = static_cast<int64_t>(part) - 1;
            uint32_t spins = 0;
            while (p >= 0) {
                const uint64_t s = __hip_atomic_load(status + p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                const uint64_t flag = s & ~kValueMask;
                if (flag == 0) {
                    // A predecessor that has its ticket is running; it publishes
                    // within microseconds. The bound only guards a broken launch.
                    if (++spins > kSpinLimit) {
                        atomicOr(error, 1u);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                prefix += static_cast<uint32_t>(s & kValueMask);
                if (flag == kInclusive) break;
                --p;
            }
            __hip_atomic_store(mine, kInclusive | static_cast<uint32_t>(prefix + blockTotal), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        }
        sPrefix = prefix;
        if (static_cast<uint64_t>(part + 1) * kPart >= n) *total = prefix + blockTotal;
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

}   // namespace

void fill32(uint32_t* dst, uint32_t value, uint64_t count, hipStream_t stream) {
    if (count == 0) return;
    const uint64_t blocks = (count + 255) / 256;
    const uint32_t grid = static_cast<uint32_t>(blocks < 4096 ? blocks : 4096);
    fill32Kernel<<<grid, 256, 0, stream>>>(dst, value, count);
}

uint32_t scanPartitions(uint32_t n) { return (n + kPart - 1) / kPart; }

void scanExclusive(const uint32_t* in, uint32_t* out, uint32_t n, uint64_t* status, uint32_t* ticket,
                   uint32_t* total, uint32_t* nonzero, uint32_t* error, hipStream_t stream) {
    const uint32_t parts = scanPartitions(n);
    fill32(reinterpret_cast<uint32_t*>(status), 0u, static_cast<uint64_t>(parts) * 2, stream);
    fill32(ticket, 0u, 1, stream);
    fill32(nonzero, 0u, 1, stream);
    if (n == 0) {
        fill32(total, 0u, 1, stream);
        return;
    }
    // 16-byte loads and stores only where both arrays allow them.
    const bool vectorised = (reinterpret_cast<uintptr_t>(in) % 16 == 0) && (reinterpret_cast<uintptr_t>(out) % 16 == 0);
    scanKernel<<<parts, kBlock, 0, stream>>>(in, out, n, status, ticket, total, nonzero, error, vectorised);
}

}   // namespace lrt::kern
