// Stable LSD radix sort of (u64 key, u32 value) pairs: one histogram pass over
// the keys for every digit at once, then one onesweep launch per 8-bit digit.
//
// A onesweep workgroup (256 threads, 4 wave64s) takes partition blockIdx.x of
// 5120 pairs and
//   1. ranks its keys by digit within each wave with ballot-built match masks
//      (eight __ballot per 64 keys), keeping a per-wave digit histogram in LDS;
//   2. turns the four wave histograms into per-wave offsets and a block digit
//      count, publishes the count for its partition and looks back over the
//      earlier partitions' published counts/prefixes (decoupled look-back) for
//      the number of equal digits before it;
//   3. scatters its pairs into LDS in local digit order and writes each digit's
//      run to global memory contiguously.
// Ranking in (wave, key slot, lane) order, which is input order, and
// partitions numbered in input order make every pass stable, so the sort is.
//
// Partitions are numbered by workgroup id, not by an atomic ticket: one
// same-address device atomic costs tens of nanoseconds across the eight XCDs,
// and a ticket per partition made that the floor of a pass. Workgroups are
// dispatched to each XCD in ascending id order, so the lowest unfinished
// partition is always resident and every look-back ends; the spin bound only
// guards a broken launch. The digit histogram likewise writes one row per
// workgroup, summed by the offsets kernel, rather than adding into one global
// row with atomics.
//
// Replaces the reference's chunk-serial LSD radix (modules/gpu/src/algo/
// RadixSort.cpp:80-158, shaders/lrt/algo/radix_*.slang) with the same result:
// a stable sort on the chosen key bits.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "kernels.h"

namespace lrt::kern {
namespace {

constexpr uint32_t kBlock = 256;
constexpr uint32_t kWaves = kBlock / 64;
constexpr uint32_t kKpt = 20;
// 5120 pairs a partition: on a 694M-pair sort 20 a thread beat 12, 15, 18,
// 24 and 30 (longer digit runs and fewer look-backs against the occupancy
// that registers and LDS allow).
constexpr uint32_t kPart = kBlock * kKpt;
constexpr uint32_t kRadix = 256;
constexpr uint32_t kMaxPasses = 8;
constexpr uint32_t kAggregate = 1u << 30;
constexpr uint32_t kInclusive = 2u << 30;
constexpr uint32_t kValueMask = (1u << 30) - 1;
constexpr uint32_t kSpinLimit = 1u << 18;
constexpr uint32_t kLookBack = 8;    // predecessor partitions read a look-back step
constexpr uint32_t kHistBlock = 1024;
constexpr uint32_t kHistRows = 512;     // histogram workgroups at most (one row of counts each)
constexpr uint32_t kHistKeys = 8192;    // keys a histogram workgroup below that cap

// The pairs to sort: the host's count, or the one a kernel left in device
// memory (then the host's is the capacity and bounds it).
__device__ __forceinline__ uint32_t pairCount(uint32_t n, const uint32_t* __restrict__ nDevice) {
    if (nDevice == nullptr) return n;
    const uint32_t d = *nDevice;
    return d < n ? d : n;
}

// TileDepth's virtual key: the tile id over the depth key less the frame's
// least (a key outside the layout's range is the duplication's error, never sorted).
__device__ __forceinline__ uint64_t tileDepthKey(uint64_t key, uint32_t depthMin, uint32_t depthBits) {
    const uint32_t depth = ((static_cast<uint32_t>(key) & 0x7FFFFFFFu) >> 7) - depthMin;
    return ((key >> 32) << depthBits) | depth;
}

template <KeyMode MODE>
__device__ __forceinline__ uint64_t sortKey(uint64_t key, uint32_t bitLo, uint64_t mask, TileDepthLayout depth) {
    if constexpr (MODE == KeyMode::TileDepth) {
        return tileDepthKey(key, depth.depthMin, depth.depthBits) & mask;
    } else {
        return (key >> bitLo) & mask;
    }
}

// Digit counts of every pass at once: workgroup b writes its row
// hist[b * passes * 256 + pass * 256 + digit]. The 64 lanes of a wave hold 64
// consecutive keys, and consecutive keys often share a digit (a splat's pairs
// share its depth; neighbouring tiles share the tile id's high bits), so each
// run of equal digits among neighbouring lanes is added as two terms: minus
// its first lane by that lane, and its last lane plus one by that lane. A run
// costs two LDS atomics whatever its length, and no lane searches for the
// run's end.
template <KeyMode MODE>
__global__ __launch_bounds__(kHistBlock) void histogramKernel(const uint64_t* __restrict__ keys, uint32_t cap,
                                                              const uint32_t* __restrict__ nDevice, uint32_t bitLo,
                                                              uint64_t mask, TileDepthLayout depth, uint32_t passes,
                                                              uint32_t* __restrict__ hist,
                                                              uint32_t* __restrict__ clear, uint64_t clearWords,
                                                              uint32_t* __restrict__ clear2, uint64_t clear2Words) {
    __shared__ uint32_t sHist[kMaxPasses * kRadix];
    // The passes' look-back words (and the caller's region), zeroed here rather
    // than by launches of their own.
    for (uint64_t w = uint64_t{blockIdx.x} * kHistBlock + threadIdx.x; w < clearWords + clear2Words;
         w += uint64_t{gridDim.x} * kHistBlock) {
        if (w < clearWords) {
            clear[w] = 0u;
        } else {
            clear2[w - clearWords] = 0u;
        }
    }
    const uint32_t n = pairCount(cap, nDevice);
    for (uint32_t j = threadIdx.x; j < passes * kRadix; j += kHistBlock) sHist[j] = 0;
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t stride = gridDim.x * kHistBlock;
    // The loop bound is the wave's first index, so whole waves iterate together.
    const uint32_t waveStart = blockIdx.x * kHistBlock + (threadIdx.x & ~63u);
    for (uint32_t w0 = waveStart; w0 < n; w0 += 4 * stride) {
        uint64_t k[4];
#pragma unroll
        for (uint32_t u = 0; u < 4; ++u) {
            const uint32_t i = w0 + u * stride + lane;   // n < 2^30: no wrap
            k[u] = i < n ? sortKey<MODE>(keys[i], bitLo, mask, depth) : 0ull;
        }
#pragma unroll
        for (uint32_t u = 0; u < 4; ++u) {
            const uint32_t first = w0 + u * stride;
            if (first >= n) break;   // uniform across the wave
            const uint32_t live = n - first < 64u ? n - first : 64u;   // lanes [0, live) hold keys
            const bool in = lane < live;
            const bool lastLive = lane + 1 == live;
            for (uint32_t p = 0; p < passes; ++p) {
                const uint32_t dg = static_cast<uint32_t>((k[u] >> (8 * p)) & 255u);
                // The digits of the lanes below and above (wave_shr:1 and
                // wave_shl:1); lanes 0 and 63 read a value unequal to their own.
                const int other = static_cast<int>(dg ^ 1u);
                const uint32_t prev =
                    static_cast<uint32_t>(__builtin_amdgcn_update_dpp(other, static_cast<int>(dg), 0x138, 0xF, 0xF, false));
                const uint32_t next =
                    static_cast<uint32_t>(__builtin_amdgcn_update_dpp(other, static_cast<int>(dg), 0x130, 0xF, 0xF, false));
                uint32_t* bin = &sHist[p * kRadix + dg];
                if (in && prev != dg) atomicAdd(bin, 0u - lane);
                if (in && (next != dg || lastLive)) atomicAdd(bin, lane + 1u);
            }
        }
    }
    __syncthreads();
    uint32_t* row = hist + static_cast<uint64_t>(blockIdx.x) * passes * kRadix;
    for (uint32_t j = threadIdx.x; j < passes * kRadix; j += kHistBlock) row[j] = sHist[j];
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

// One workgroup per pass: the histogram rows summed (four row groups of 256
// threads, eight rows in flight a thread), then the exclusive scan of the 256
// digit counts.
__global__ __launch_bounds__(1024) void digitOffsetsKernel(const uint32_t* __restrict__ hist, uint32_t rows,
                                                           uint32_t passes, uint32_t* __restrict__ offsets) {
    __shared__ uint32_t sGroup[4][kRadix];
    __shared__ uint32_t sWave[kRadix / 64];
    const uint32_t t = threadIdx.x, digit = t & 255u, group = t >> 8;
    const uint64_t rowWords = static_cast<uint64_t>(passes) * kRadix;
    const uint32_t* col = hist + static_cast<uint64_t>(blockIdx.x) * kRadix + digit;
    uint32_t sum = 0;
    for (uint32_t r0 = group; r0 < rows; r0 += 32) {
        uint32_t part[8];
#pragma unroll
        for (uint32_t u = 0; u < 8; ++u) {
            const uint32_t r = r0 + 4 * u;
            part[u] = r < rows ? col[r * rowWords] : 0u;
        }
#pragma unroll
        for (uint32_t u = 0; u < 8; ++u) sum += part[u];
    }
    sGroup[group][digit] = sum;
    __syncthreads();
    // Threads 0-255 (whole waves 0-3) scan; every thread reaches each barrier.
    const uint32_t lane = t & 63u, wave = t >> 6;
    uint32_t c = 0, incl = 0;
    if (t < kRadix) {
        c = sGroup[0][t] + sGroup[1][t] + sGroup[2][t] + sGroup[3][t];
        incl = waveInclusiveScan(c, lane);
        if (lane == 63) sWave[wave] = incl;
    }
    __syncthreads();
    if (t < kRadix) {
        uint32_t base = 0;
        for (uint32_t w = 0; w < wave; ++w) base += sWave[w];
        offsets[blockIdx.x * kRadix + t] = base + incl - c;
    }
}

// How a pass reads its keys, takes its 8-bit digit and writes its keys:
//   Plain64:      u64 in and out; the digit is bits [shift, +8) of the key.
//   TileDepth64:  u64 in and out, the duplication's keys; the digit is bits
//                 [shift, +8) of the virtual key (tileDepthKey).
//   TileNarrow:   u64 in (the duplication's keys), u32 out: the virtual key
//                 shifted right by `narrow` (its low `narrow` bits are this
//                 pass's, consumed once it is done); the digit as TileDepth64's.
//   Narrow32:     u32 in and out (TileNarrow's output); the digit is bits
//                 [shift, +8) of the stored key.
// `mask` keeps the digit's bits below keyBits.
enum class PassKind : uint32_t { Plain64, TileDepth64, TileNarrow, Narrow32 };

struct PassDigit {
    uint32_t shift;
    uint32_t mask;
    uint32_t depthMin, depthBits;   // TileDepth's layout
    uint32_t narrow;                // TileNarrow: the bits dropped from the virtual key
};

template <PassKind K>
struct PassKeys {
    using In = std::conditional_t<K == PassKind::Narrow32, uint32_t, uint64_t>;
    using Out = std::conditional_t<K == PassKind::TileNarrow || K == PassKind::Narrow32, uint32_t, uint64_t>;
};

// The digit of an input key and the key as the pass writes it.
template <PassKind K>
__device__ __forceinline__ uint32_t digitOf(typename PassKeys<K>::In key, const PassDigit& pd,
                                            typename PassKeys<K>::Out& out) {
    if constexpr (K == PassKind::TileDepth64 || K == PassKind::TileNarrow) {
        const uint64_t v = tileDepthKey(key, pd.depthMin, pd.depthBits);
        if constexpr (K == PassKind::TileNarrow) {
            out = static_cast<uint32_t>(v >> pd.narrow);
        } else {
            out = key;
        }
        return static_cast<uint32_t>(v >> pd.shift) & pd.mask;
    } else {
        out = key;
        return static_cast<uint32_t>(key >> pd.shift) & pd.mask;
    }
}

template <PassKind K>
__global__ __launch_bounds__(kBlock) void onesweepKernel(const typename PassKeys<K>::In* __restrict__ keysIn,
                                                         const uint32_t* __restrict__ valsIn,
                                                         typename PassKeys<K>::Out* __restrict__ keysOut,
                                                         uint32_t* __restrict__ valsOut, uint32_t cap,
                                                         const uint32_t* __restrict__ nDevice, PassDigit pd,
                                                         const uint32_t* __restrict__ digitOffsets,
                                                         uint32_t* __restrict__ status,
                                                         uint32_t* __restrict__ error) {
    using KIn = typename PassKeys<K>::In;
    using KOut = typename PassKeys<K>::Out;
    __shared__ KOut sKeys[kPart];
    __shared__ uint32_t sVals[kPart];
    // Per-wave digit counts; once the scatter positions are in registers, the
    // digits of sKeys (bytes), so the write-out need not recompute them.
    constexpr uint32_t kCountWords = kWaves * kRadix > (kPart + 3) / 4 ? kWaves * kRadix : (kPart + 3) / 4;
    __shared__ uint32_t sCounts[kCountWords];
    __shared__ uint32_t sLocal[kRadix];
    __shared__ uint32_t sGlobal[kRadix];
    __shared__ uint32_t sScan[kWaves];
    auto sWave = reinterpret_cast<uint32_t (*)[kRadix]>(sCounts);
    uint8_t* sDigit = reinterpret_cast<uint8_t*>(sCounts);

    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u;
    const uint32_t wave = tid >> 6;
#pragma unroll
    for (uint32_t w = 0; w < kWaves; ++w) sWave[w][tid] = 0;
    __syncthreads();
    const uint32_t n = pairCount(cap, nDevice);
    const uint32_t part = blockIdx.x;
    const uint32_t base = part * kPart;
    // Launched for the capacity: partitions past the count have nothing to
    // do, and no partition before them waits on them.
    if (base >= n) return;
    const uint32_t waveFirst = base + wave * (kKpt * 64);

    KIn kin[kKpt];
    KOut k[kKpt];
    uint32_t v[kKpt];
    uint32_t d[kKpt];
    uint32_t r[kKpt];
    // Loads clamped to the last pair (no branches); lanes past n are masked
    // out of every vote below and write nothing.
    const uint32_t last = n - 1;
#pragma unroll
    for (uint32_t j = 0; j < kKpt; ++j) {
        const uint32_t idx = min(waveFirst + j * 64 + lane, last);
        kin[j] = keysIn[idx];
        v[j] = valsIn[idx];
    }
#pragma unroll
    for (uint32_t j = 0; j < kKpt; ++j) d[j] = digitOf<K>(kin[j], pd, k[j]);

    // Rank within the wave: peers = lanes holding the same digit, the AND over
    // the digit's bits of the bit's vote (bit set) or its complement (clear);
    // m = all ones where the bit is clear, so vote ^ m is that term.
#pragma unroll
    for (uint32_t j = 0; j < kKpt; ++j) {
        const bool valid = waveFirst + j * 64 + lane < n;
        const uint64_t live = __ballot(valid);
        uint32_t lo = static_cast<uint32_t>(live), hi = static_cast<uint32_t>(live >> 32);
        const int nd = static_cast<int>(~d[j]);
#pragma unroll
        for (uint32_t b = 0; b < 8; ++b) {
            const uint32_t m = static_cast<uint32_t>(__builtin_amdgcn_sbfe(nd, b, 1));
            const uint64_t vote = __ballot(m == 0u);
            lo &= static_cast<uint32_t>(vote) ^ m;
            hi &= static_cast<uint32_t>(vote >> 32) ^ m;
        }
        const uint32_t before = __builtin_amdgcn_mbcnt_hi(hi, __builtin_amdgcn_mbcnt_lo(lo, 0u));
        const uint32_t count = static_cast<uint32_t>(__popc(lo) + __popc(hi));
        // The count a leader stores is read by other lanes of the wave at the
        // next key: relaxed atomics, so the compiler keeps no per-lane copy.
        // Every lane reads (a lane past n holds the digit of a clamped load),
        // so no branch holds the read back.
        const uint32_t prev = __hip_atomic_load(&sWave[wave][d[j]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
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
        // Look back kLookBack partitions a step: the window's loads are in
        // flight together, so a walk of L partitions costs about L / kLookBack
        // round trips to L2 instead of L. Eight measured best on sorts of 2M
        // and of 694M pairs: a wider window's status reads (1 KiB a
        // partition) cost more than the round trips it saves. Partition 0 is
        // inclusive from the start, so no walk reaches below it.
        static_assert(kLookBack < 32, "a window's flags fit one word");
        int32_t p = static_cast<int32_t>(part) - 1;   // partitions < 2^31
        const uint32_t* column = status + tid;         // digit tid's word of partition 0
        uint32_t spins = 0;
        while (true) {
            // Words below partition 0 read partition 0's (every load issued,
            // no branch) and count as inclusive zeros.
            uint32_t s[kLookBack];
#pragma unroll
            for (uint32_t j = 0; j < kLookBack; ++j) {
                const int32_t q = p - static_cast<int32_t>(j);
                const uint32_t at = static_cast<uint32_t>(q > 0 ? q : 0) * kRadix;
                const uint32_t word = __hip_atomic_load(column + at, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                s[j] = q >= 0 ? word : kInclusive;
            }
            // The window is summed from its top down to its first unpublished
            // word, or through its first inclusive word if that comes first
            // (bit j of each mask is word j's). Flags are 1 (aggregate) or 2
            // (inclusive), never 3.
            uint32_t published = 0, inclusive = 0;
#pragma unroll
            for (uint32_t j = 0; j < kLookBack; ++j) {
                published |= (s[j] >= kAggregate ? 1u : 0u) << j;
                inclusive |= (s[j] >> 31) << j;
            }
            const uint32_t unpublished = static_cast<uint32_t>(__builtin_ctz(~published));   // kLookBack if none
            const uint32_t firstInclusive = static_cast<uint32_t>(__builtin_ctz(inclusive | (1u << kLookBack)));
            const bool done = firstInclusive < unpublished;
            const uint32_t taken = done ? firstInclusive + 1 : unpublished;
#pragma unroll
            for (uint32_t j = 0; j < kLookBack; ++j) prefix += j < taken ? s[j] & kValueMask : 0u;
            if (done) break;
            p -= static_cast<int32_t>(taken);
            if (taken == 0) {
                // The partition waited on has a lower workgroup id, so it is
                // running or next on its XCD; the bound only guards a broken launch.
                if (++spins > kSpinLimit) {
                    atomicOr(error, 1u);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
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

    // Local scatter into digit order: positions first (the wave counts are
    // read), then the pairs and their digits into LDS, then contiguous runs out.
#pragma unroll
    for (uint32_t j = 0; j < kKpt; ++j) r[j] += sLocal[d[j]] + sWave[wave][d[j]];
    __syncthreads();
    bool misplaced = false;
#pragma unroll
    for (uint32_t j = 0; j < kKpt; ++j) {
        const bool valid = waveFirst + j * 64 + lane < n;
        const uint32_t at = r[j];
        if (valid & (at < kPart)) {
            sKeys[at] = k[j];
            sVals[at] = v[j];
            sDigit[at] = static_cast<uint8_t>(d[j]);
        }
        misplaced |= valid & (at >= kPart);
    }
    if (misplaced) atomicOr(error, 2u);
    __syncthreads();
    // Every place's LDS reads are issued before any store, unconditionally
    // (places past the count read stale entries of the partition's own LDS,
    // digits below 256, and store nothing); only the stores are predicated,
    // and a bad position is flagged once after the loop.
    const uint32_t count = n - base < kPart ? n - base : kPart;
    uint32_t out[kKpt];
#pragma unroll
    for (uint32_t j = 0; j < kKpt; ++j) out[j] = sDigit[tid + j * kBlock];
#pragma unroll
    for (uint32_t j = 0; j < kKpt; ++j) out[j] = sGlobal[out[j]] + tid + j * kBlock;
#pragma unroll
    for (uint32_t j = 0; j < kKpt; ++j) {
        k[j] = sKeys[tid + j * kBlock];
        v[j] = sVals[tid + j * kBlock];
    }
    bool bad = false;
#pragma unroll
    for (uint32_t j = 0; j < kKpt; ++j) {
        const bool inside = tid + j * kBlock < count;
        if (inside & (out[j] < n)) {
            keysOut[out[j]] = k[j];
            valsOut[out[j]] = v[j];
        }
        bad |= inside & (out[j] >= n);
    }
    if (bad) atomicOr(error, 4u);
}

// Histogram workgroups for capacity n: one per kHistKeys keys, at most kHistRows.
uint32_t histogramRows(uint32_t n) {
    const uint32_t rows = (n + kHistKeys - 1) / kHistKeys;
    return std::clamp(rows, 1u, kHistRows);
}

template <KeyMode MODE>
void launchPasses(uint64_t* keys, uint32_t* values, uint64_t* keysAlt, uint32_t* valuesAlt, uint32_t n,
                  const uint32_t* nDevice, uint32_t bitLo, uint32_t keyBits, TileDepthLayout depth,
                  const SortWorkspace& ws, hipStream_t stream, hipEvent_t afterHistogram, uint32_t* alsoZero,
                  uint64_t alsoZeroWords) {
    const uint32_t passes = sortPasses(keyBits);
    const uint32_t parts = sortPartitions(n);
    const uint64_t mask = keyBits >= 64 ? ~0ull : ((1ull << keyBits) - 1ull);
    // Without a device count the histogram kernel zeroes the look-back words.
    const uint64_t clearWords = nDevice == nullptr ? static_cast<uint64_t>(passes) * parts * kRadix : 0u;
    if (nDevice != nullptr) {
        const uint64_t words = static_cast<uint64_t>(passes) * parts * kRadix;
        const uint32_t grid = static_cast<uint32_t>(std::min<uint64_t>((words + kBlock - 1) / kBlock, 4096));
        clearStatusKernel<<<grid, kBlock, 0, stream>>>(ws.status, passes, parts, n, nDevice);
    }
    const uint32_t rows = histogramRows(n);
    histogramKernel<MODE><<<rows, kHistBlock, 0, stream>>>(keys, n, nDevice, bitLo, mask, depth, passes,
                                                          ws.histogram, ws.status, clearWords, alsoZero,
                                                          alsoZero != nullptr ? alsoZeroWords : 0u);
    digitOffsetsKernel<<<passes, 1024, 0, stream>>>(ws.histogram, rows, passes, ws.offsets);
    if (afterHistogram != nullptr) (void)hipEventRecord(afterHistogram, stream);
    // TileDepth keys of at most kNarrowMaxBits bits leave the first pass as
    // u32 (the virtual key less its low narrowShift bits, which that pass
    // consumes), so the later passes move 8 bytes a pair instead of 12.
    const bool narrow = MODE == KeyMode::TileDepth && keyBits <= kNarrowMaxBits;
    const uint32_t dropped = narrowShift(keyBits);
    const void* kin = keys;
    const uint32_t* vin = values;
    void* kout = keysAlt;
    uint32_t* vout = valuesAlt;
    for (uint32_t p = 0; p < passes; ++p) {
        PassDigit pd{8 * p, static_cast<uint32_t>((mask >> (8 * p)) & 255u), depth.depthMin, depth.depthBits, dropped};
        const uint32_t* offsets = ws.offsets + p * kRadix;
        uint32_t* status = ws.status + static_cast<uint64_t>(p) * parts * kRadix;
        if constexpr (MODE == KeyMode::Plain) {
            pd.shift = bitLo + 8 * p;
            onesweepKernel<PassKind::Plain64><<<parts, kBlock, 0, stream>>>(
                static_cast<const uint64_t*>(kin), vin, static_cast<uint64_t*>(kout), vout, n, nDevice, pd, offsets,
                status, ws.error);
        } else if (!narrow) {
            onesweepKernel<PassKind::TileDepth64><<<parts, kBlock, 0, stream>>>(
                static_cast<const uint64_t*>(kin), vin, static_cast<uint64_t*>(kout), vout, n, nDevice, pd, offsets,
                status, ws.error);
        } else if (p == 0) {
            onesweepKernel<PassKind::TileNarrow><<<parts, kBlock, 0, stream>>>(
                static_cast<const uint64_t*>(kin), vin, static_cast<uint32_t*>(kout), vout, n, nDevice, pd, offsets,
                status, ws.error);
        } else {
            pd.shift = 8 * p - dropped;
            onesweepKernel<PassKind::Narrow32><<<parts, kBlock, 0, stream>>>(
                static_cast<const uint32_t*>(kin), vin, static_cast<uint32_t*>(kout), vout, n, nDevice, pd, offsets,
                status, ws.error);
        }
        const void* nk = kout;
        const uint32_t* nv = vout;
        kout = const_cast<void*>(kin);
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
    return uint64_t{histogramRows(n)} * passes * kRadix + passes * kRadix + passes * sortPartitions(n) * kRadix;
}

SortWorkspace sortWorkspace(uint32_t* work, uint32_t n, uint32_t keyBits, uint32_t* error) {
    const uint64_t passes = sortPasses(keyBits);
    const uint64_t histWords = uint64_t{histogramRows(n)} * passes * kRadix;
    return SortWorkspace{work, work + histWords, work + histWords + passes * kRadix, error};
}

bool sortPairs(uint64_t* keys, uint32_t* values, uint64_t* keysAlt, uint32_t* valuesAlt, uint32_t n, KeyMode mode,
               uint32_t bitLo, uint32_t keyBits, const SortWorkspace& ws, hipStream_t stream,
               hipEvent_t afterHistogram, const uint32_t* nDevice, TileDepthLayout depth, uint32_t* alsoZero,
               uint64_t alsoZeroWords) {
    if ((nDevice == nullptr && n < 2) || n == 0 || keyBits == 0) {
        if (alsoZero != nullptr) fill32(alsoZero, 0u, alsoZeroWords, stream);
        if (afterHistogram != nullptr) (void)hipEventRecord(afterHistogram, stream);
        return false;
    }
    if (mode == KeyMode::TileDepth) {
        launchPasses<KeyMode::TileDepth>(keys, values, keysAlt, valuesAlt, n, nDevice, bitLo, keyBits, depth, ws,
                                         stream, afterHistogram, alsoZero, alsoZeroWords);
    } else {
        launchPasses<KeyMode::Plain>(keys, values, keysAlt, valuesAlt, n, nDevice, bitLo, keyBits, depth, ws, stream,
                                     afterHistogram, alsoZero, alsoZeroWords);
    }
    return (sortPasses(keyBits) & 1u) != 0u;
}

}   // namespace lrt::kern
