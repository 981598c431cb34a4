// Stable LSD radix sort of (u64 key, u32 value) pairs: one histogram pass over
// the keys for every digit at once, then one onesweep launch per digit. A
// digit is 8 bits or fewer, or up to 10 where that saves a pass (sortPlan).
//
// A onesweep workgroup (256 threads, 4 wave64s) takes partition blockIdx.x of
// 5120 pairs and
//   1. ranks its keys by digit within each wave with ballot-built match masks
//      (one __ballot per digit bit per 64 keys), keeping a per-wave digit
//      histogram in LDS;
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
constexpr uint32_t kMinTableBits = 8;    // digit tables hold 2^max(8, width) entries
constexpr uint32_t kMaxRadixBits = 10;
constexpr uint32_t kAggregate = 1u << 30;
constexpr uint32_t kInclusive = 2u << 30;
constexpr uint32_t kValueMask = (1u << 30) - 1;
constexpr uint32_t kSpinLimit = 1u << 18;
constexpr uint32_t kLookBack = 8;    // predecessor partitions read a look-back step (8-bit digits)
constexpr uint32_t kHistBlock = 1024;
constexpr uint32_t kHistRows = 512;     // histogram workgroups at most (one row of counts each)
constexpr uint32_t kHistKeys = 8192;    // keys a histogram workgroup below that cap
constexpr uint32_t kHistBins = 6144;    // a plan's digit tables, all passes (60 bits at radix 10 the most)

// A plan's digits as the kernels take them: pass p's digit is bits
// [shift[p], +width) of the key under mask[p]; its tables hold 2^tableBits[p]
// entries at bins[p] of a row of totalBins.
struct HistPlan {
    uint32_t passes, totalBins;
    uint32_t shift[kMaxSortPasses], mask[kMaxSortPasses], bins[kMaxSortPasses], tableBits[kMaxSortPasses];
};

HistPlan histPlanOf(const SortPlan& plan) {
    HistPlan h{};
    h.passes = plan.passes;
    h.totalBins = plan.totalBins;
    for (uint32_t p = 0; p < plan.passes; ++p) {
        h.shift[p] = plan.shift[p];
        h.mask[p] = (1u << plan.width[p]) - 1u;
        h.bins[p] = plan.bins[p];
        h.tableBits[p] = std::max(plan.width[p], kMinTableBits);
    }
    return h;
}

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
// hist[b * totalBins + bins[pass] + digit]. The 64 lanes of a wave hold 64
// consecutive keys, and consecutive keys often share a digit (a splat's pairs
// share its depth; neighbouring tiles share the tile id's high bits), so each
// run of equal digits among neighbouring lanes is added as two terms: minus
// its first lane by that lane, and its last lane plus one by that lane. A run
// costs two LDS atomics whatever its length, and no lane searches for the
// run's end.
template <KeyMode MODE>
__global__ __launch_bounds__(kHistBlock) void histogramKernel(const uint64_t* __restrict__ keys, uint32_t cap,
                                                              const uint32_t* __restrict__ nDevice, uint32_t bitLo,
                                                              uint64_t mask, TileDepthLayout depth, HistPlan plan,
                                                              uint32_t* __restrict__ hist,
                                                              uint32_t* __restrict__ clear, uint64_t clearWords,
                                                              uint32_t* __restrict__ clear2, uint64_t clear2Words) {
    __shared__ uint32_t sHist[kHistBins];
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
    for (uint32_t j = threadIdx.x; j < plan.totalBins; j += kHistBlock) sHist[j] = 0;
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
            for (uint32_t p = 0; p < plan.passes; ++p) {
                const uint32_t dg = static_cast<uint32_t>(k[u] >> plan.shift[p]) & plan.mask[p];
                // The digits of the lanes below and above (wave_shr:1 and
                // wave_shl:1); lanes 0 and 63 read a value unequal to their own.
                const int other = static_cast<int>(dg ^ 1u);
                const uint32_t prev =
                    static_cast<uint32_t>(__builtin_amdgcn_update_dpp(other, static_cast<int>(dg), 0x138, 0xF, 0xF, false));
                const uint32_t next =
                    static_cast<uint32_t>(__builtin_amdgcn_update_dpp(other, static_cast<int>(dg), 0x130, 0xF, 0xF, false));
                uint32_t* bin = &sHist[plan.bins[p] + dg];
                if (in && prev != dg) atomicAdd(bin, 0u - lane);
                if (in && (next != dg || lastLive)) atomicAdd(bin, lane + 1u);
            }
        }
    }
    __syncthreads();
    uint32_t* row = hist + static_cast<uint64_t>(blockIdx.x) * plan.totalBins;
    for (uint32_t j = threadIdx.x; j < plan.totalBins; j += kHistBlock) row[j] = sHist[j];
}

// Clears the look-back words of the partitions the count needs, every pass
// (pass p's words of partition q at capParts * bins[p] + q * 2^tableBits[p]).
__global__ __launch_bounds__(kBlock) void clearStatusKernel(uint32_t* __restrict__ status, HistPlan plan,
                                                            uint32_t capParts, uint32_t cap,
                                                            const uint32_t* __restrict__ nDevice) {
    const uint32_t n = pairCount(cap, nDevice);
    const uint64_t used = (n + kPart - 1) / kPart;
    const uint64_t total = used * plan.totalBins;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; i < total; i += stride) {
        // Pass p's used words are [used * bins[p], used * bins[p + 1]) of i's range.
        uint32_t p = 0;
        while (p + 1 < plan.passes && i >= used * plan.bins[p + 1]) ++p;
        status[static_cast<uint64_t>(capParts) * plan.bins[p] + (i - used * plan.bins[p])] = 0u;
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

// One workgroup per pass: the histogram rows summed (1024 / R row groups of R
// threads for the pass's R = 2^tableBits digits, eight rows in flight a
// thread), then the exclusive scan of the R digit counts.
__global__ __launch_bounds__(1024) void digitOffsetsKernel(const uint32_t* __restrict__ hist, uint32_t rows,
                                                           HistPlan plan, uint32_t* __restrict__ offsets) {
    __shared__ uint32_t sGroup[1024];
    __shared__ uint32_t sWave[1024 / 64];
    const uint32_t pass = blockIdx.x;
    const uint32_t tableBits = plan.tableBits[pass], radix = 1u << tableBits, groups = 1024u >> tableBits;
    const uint32_t t = threadIdx.x, digit = t & (radix - 1u), group = t >> tableBits;
    const uint32_t* col = hist + plan.bins[pass] + digit;
    uint32_t sum = 0;
    for (uint32_t r0 = group; r0 < rows; r0 += 8 * groups) {
        uint32_t part[8];
#pragma unroll
        for (uint32_t u = 0; u < 8; ++u) {
            const uint32_t r = r0 + groups * u;
            part[u] = r < rows ? col[static_cast<uint64_t>(r) * plan.totalBins] : 0u;
        }
#pragma unroll
        for (uint32_t u = 0; u < 8; ++u) sum += part[u];
    }
    sGroup[t] = sum;   // sGroup[group * radix + digit]
    __syncthreads();
    // Threads below radix (whole waves) scan; every thread reaches each barrier.
    const uint32_t lane = t & 63u, wave = t >> 6;
    uint32_t c = 0, incl = 0;
    if (t < radix) {
        for (uint32_t g = 0; g < groups; ++g) c += sGroup[g * radix + t];
        incl = waveInclusiveScan(c, lane);
        if (lane == 63) sWave[wave] = incl;
    }
    __syncthreads();
    if (t < radix) {
        uint32_t base = 0;
        for (uint32_t w = 0; w < wave; ++w) base += sWave[w];
        offsets[plan.bins[pass] + t] = base + incl - c;
    }
}

// How a pass reads its keys, takes its digit and writes its keys:
//   Plain64:      u64 in and out; the digit is bits [shift, +width) of the key.
//   TileDepth64:  u64 in and out, the duplication's keys; the digit is bits
//                 [shift, +width) of the virtual key (tileDepthKey).
//   TileNarrow:   u64 in (the duplication's keys), u32 out: the virtual key
//                 shifted right by `narrow` (its low `narrow` bits are this
//                 pass's, consumed once it is done); the digit as TileDepth64's.
//   Narrow32:     u32 in and out (TileNarrow's output); the digit is bits
//                 [shift, +width) of the stored key.
// `mask` holds the digit's width in ones.
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

template <PassKind K, uint32_t W>
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
    static_assert(W >= kMinTableBits && W <= kMaxRadixBits, "digit table size");
    constexpr uint32_t kR = 1u << W;            // digits in the tables
    constexpr uint32_t kD = kR / kBlock;        // digits a thread looks back for
    // Look-back window: 8 partitions at 256 digits, fewer at wider digits, so a
    // thread has eight status loads in flight a step (two at least).
    constexpr uint32_t kLB = (kLookBack * kBlock / kR) > 2 ? (kLookBack * kBlock / kR) : 2;
    using DigitT = std::conditional_t<(W > 8), uint16_t, uint8_t>;
    __shared__ KOut sKeys[kPart];
    __shared__ uint32_t sVals[kPart];
    // Per-wave digit counts, then each (wave, digit)'s first place in the
    // partition; once the scatter positions are in registers, the digits of
    // sKeys, so the write-out need not recompute them.
    constexpr uint32_t kDigitWords = (kPart * static_cast<uint32_t>(sizeof(DigitT)) + 3) / 4;
    constexpr uint32_t kCountWords = kWaves * kR > kDigitWords ? kWaves * kR : kDigitWords;
    __shared__ uint32_t sCounts[kCountWords];
    __shared__ uint32_t sGlobal[kR];
    __shared__ uint32_t sScan[kD][kWaves];
    auto sWave = reinterpret_cast<uint32_t (*)[kR]>(sCounts);
    DigitT* sDigit = reinterpret_cast<DigitT*>(sCounts);

    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u;
    const uint32_t wave = tid >> 6;
#pragma unroll
    for (uint32_t w = 0; w < kWaves; ++w) {
#pragma unroll
        for (uint32_t i = 0; i < kD; ++i) sWave[w][i * kBlock + tid] = 0;
    }
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
    // m = all ones where the bit is clear, so vote ^ m is that term. Bits past
    // the pass's width are clear in every digit, their terms all ones.
#pragma unroll
    for (uint32_t j = 0; j < kKpt; ++j) {
        const bool valid = waveFirst + j * 64 + lane < n;
        const uint64_t live = __ballot(valid);
        uint32_t lo = static_cast<uint32_t>(live), hi = static_cast<uint32_t>(live >> 32);
        const int nd = static_cast<int>(~d[j]);
#pragma unroll
        for (uint32_t b = 0; b < W; ++b) {
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

    // Digits i * kBlock + tid: wave counts, block count, publication and look-back.
    uint32_t c[kD][kWaves], total[kD], incl[kD], prefix[kD];
    uint32_t* const mine = status + static_cast<uint64_t>(part) * kR + tid;
#pragma unroll
    for (uint32_t i = 0; i < kD; ++i) {
        total[i] = 0;
#pragma unroll
        for (uint32_t w = 0; w < kWaves; ++w) {
            c[i][w] = sWave[w][i * kBlock + tid];
            total[i] += c[i][w];
        }
        __hip_atomic_store(mine + i * kBlock, (part == 0 ? kInclusive : kAggregate) | total[i], __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
        incl[i] = waveInclusiveScan(total[i], lane);
        if (lane == 63) sScan[i][wave] = incl[i];
        prefix[i] = 0;
    }
    if (part > 0) {
        // Look back kLB partitions a step: the window's loads are in flight
        // together, so a walk of L partitions costs about L / kLB round trips
        // to L2 instead of L. At 256 digits eight measured best on sorts of 2M
        // and of 694M pairs: a wider window's status reads (1 KiB a partition)
        // cost more than the round trips it saves. Partition 0 is inclusive
        // from the start, so no walk reaches below it. Each digit walks on its
        // own; the loop ends when every digit of the thread has.
        static_assert(kLB < 32, "a window's flags fit one word");
        int32_t at[kD];   // the newest partition digit i has not summed yet
#pragma unroll
        for (uint32_t i = 0; i < kD; ++i) at[i] = static_cast<int32_t>(part) - 1;   // partitions < 2^31
        const uint32_t* column = status + tid;   // digit tid's word of partition 0
        uint32_t pending = (1u << kD) - 1u;
        uint32_t spins = 0;
        while (pending != 0) {
            // Words below partition 0 read partition 0's (every load issued,
            // no branch) and count as inclusive zeros.
            uint32_t sw[kD][kLB];
#pragma unroll
            for (uint32_t i = 0; i < kD; ++i) {
#pragma unroll
                for (uint32_t j = 0; j < kLB; ++j) {
                    const int32_t q = at[i] - static_cast<int32_t>(j);
                    const uint32_t word = __hip_atomic_load(
                        column + static_cast<uint64_t>(q > 0 ? q : 0) * kR + i * kBlock, __ATOMIC_RELAXED,
                        __HIP_MEMORY_SCOPE_AGENT);
                    sw[i][j] = q >= 0 ? word : kInclusive;
                }
            }
            bool stalled = false;
#pragma unroll
            for (uint32_t i = 0; i < kD; ++i) {
                if ((pending >> i) & 1u) {
                    // The window is summed from its top down to its first
                    // unpublished word, or through its first inclusive word if
                    // that comes first (bit j of each mask is word j's). Flags
                    // are 1 (aggregate) or 2 (inclusive), never 3.
                    uint32_t published = 0, inclusive = 0;
#pragma unroll
                    for (uint32_t j = 0; j < kLB; ++j) {
                        published |= (sw[i][j] >= kAggregate ? 1u : 0u) << j;
                        inclusive |= (sw[i][j] >> 31) << j;
                    }
                    const uint32_t unpublished = static_cast<uint32_t>(__builtin_ctz(~published));   // kLB if none
                    const uint32_t firstInclusive = static_cast<uint32_t>(__builtin_ctz(inclusive | (1u << kLB)));
                    const bool done = firstInclusive < unpublished;
                    const uint32_t taken = done ? firstInclusive + 1 : unpublished;
#pragma unroll
                    for (uint32_t j = 0; j < kLB; ++j) prefix[i] += j < taken ? sw[i][j] & kValueMask : 0u;
                    if (done) {
                        pending &= ~(1u << i);
                        __hip_atomic_store(mine + i * kBlock, kInclusive | (prefix[i] + total[i]), __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
                    } else {
                        at[i] -= static_cast<int32_t>(taken);
                        stalled |= taken == 0;
                    }
                }
            }
            if (stalled) {
                // The partition waited on has a lower workgroup id, so it is
                // running or next on its XCD; the bound only guards a broken launch.
                if (++spins > kSpinLimit) {
                    atomicOr(error, 1u);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
    }
    __syncthreads();
    // Each digit's first place in the partition (digits in order: chunk i of
    // kBlock digits after the chunks below it), then each (wave, digit)'s, and
    // the digit's global place less its local one.
    uint32_t chunksBelow = 0;
#pragma unroll
    for (uint32_t i = 0; i < kD; ++i) {
        uint32_t chunk = 0, waveBase = 0;
#pragma unroll
        for (uint32_t w = 0; w < kWaves; ++w) {
            const uint32_t x = sScan[i][w];
            chunk += x;
            waveBase += w < wave ? x : 0u;
        }
        const uint32_t localStart = chunksBelow + waveBase + incl[i] - total[i];
        chunksBelow += chunk;
        const uint32_t dg = i * kBlock + tid;
        uint32_t run = localStart;
#pragma unroll
        for (uint32_t w = 0; w < kWaves; ++w) {
            sWave[w][dg] = run;
            run += c[i][w];
        }
        sGlobal[dg] = digitOffsets[dg] + prefix[i] - localStart;
    }
    __syncthreads();

    // Local scatter into digit order: positions first (the wave counts are
    // read), then the pairs and their digits into LDS, then contiguous runs out.
#pragma unroll
    for (uint32_t j = 0; j < kKpt; ++j) r[j] += sWave[wave][d[j]];
    __syncthreads();
    bool misplaced = false;
#pragma unroll
    for (uint32_t j = 0; j < kKpt; ++j) {
        const bool valid = waveFirst + j * 64 + lane < n;
        const uint32_t to = r[j];
        if (valid & (to < kPart)) {
            sKeys[to] = k[j];
            sVals[to] = v[j];
            sDigit[to] = static_cast<DigitT>(d[j]);
        }
        misplaced |= valid & (to >= kPart);
    }
    if (misplaced) atomicOr(error, 2u);
    __syncthreads();
    // Every place's LDS reads are issued before any store, unconditionally
    // (places past the count read stale entries of the partition's own LDS,
    // digits below kR, and store nothing); only the stores are predicated,
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

// One onesweep launch of kind K over tables of 2^tableBits digits.
template <PassKind K>
void launchPass(uint32_t tableBits, uint32_t parts, hipStream_t stream, const void* kin, const uint32_t* vin,
                void* kout, uint32_t* vout, uint32_t n, const uint32_t* nDevice, const PassDigit& pd,
                const uint32_t* offsets, uint32_t* status, uint32_t* error) {
    using In = typename PassKeys<K>::In;
    using Out = typename PassKeys<K>::Out;
    const auto go = [&](auto w) {
        onesweepKernel<K, decltype(w)::value><<<parts, kBlock, 0, stream>>>(
            static_cast<const In*>(kin), vin, static_cast<Out*>(kout), vout, n, nDevice, pd, offsets, status, error);
    };
    if (tableBits <= 8) {
        go(std::integral_constant<uint32_t, 8>{});
    } else if (tableBits == 9) {
        go(std::integral_constant<uint32_t, 9>{});
    } else {
        go(std::integral_constant<uint32_t, 10>{});
    }
}

template <KeyMode MODE>
void launchPasses(uint64_t* keys, uint32_t* values, uint64_t* keysAlt, uint32_t* valuesAlt, uint32_t n,
                  const uint32_t* nDevice, uint32_t bitLo, uint32_t keyBits, const SortPlan& plan,
                  TileDepthLayout depth, const SortWorkspace& ws, hipStream_t stream, hipEvent_t afterHistogram,
                  uint32_t* alsoZero, uint64_t alsoZeroWords) {
    const HistPlan hp = histPlanOf(plan);
    const uint32_t parts = sortPartitions(n);
    const uint64_t mask = keyBits >= 64 ? ~0ull : ((1ull << keyBits) - 1ull);
    // Without a device count the histogram kernel zeroes the look-back words.
    const uint64_t clearWords = nDevice == nullptr ? uint64_t{parts} * plan.totalBins : 0u;
    if (nDevice != nullptr) {
        const uint64_t words = uint64_t{parts} * plan.totalBins;
        const uint32_t grid = static_cast<uint32_t>(std::min<uint64_t>((words + kBlock - 1) / kBlock, 4096));
        clearStatusKernel<<<grid, kBlock, 0, stream>>>(ws.status, hp, parts, n, nDevice);
    }
    const uint32_t rows = histogramRows(n);
    histogramKernel<MODE><<<rows, kHistBlock, 0, stream>>>(keys, n, nDevice, bitLo, mask, depth, hp, ws.histogram,
                                                          ws.status, clearWords, alsoZero,
                                                          alsoZero != nullptr ? alsoZeroWords : 0u);
    digitOffsetsKernel<<<plan.passes, 1024, 0, stream>>>(ws.histogram, rows, hp, ws.offsets);
    if (afterHistogram != nullptr) (void)hipEventRecord(afterHistogram, stream);
    // TileDepth keys at most 32 bits wider than their first digit leave the
    // first pass as u32 (the virtual key less its low narrowShift bits, which
    // that pass consumes), so the later passes move 8 bytes a pair instead of 12.
    const bool narrow = MODE == KeyMode::TileDepth && keyBits <= 32 + plan.width[0];
    const uint32_t dropped = narrowShift(keyBits);
    const void* kin = keys;
    const uint32_t* vin = values;
    void* kout = keysAlt;
    uint32_t* vout = valuesAlt;
    for (uint32_t p = 0; p < plan.passes; ++p) {
        PassDigit pd{plan.shift[p], hp.mask[p], depth.depthMin, depth.depthBits, dropped};
        const uint32_t* offsets = ws.offsets + plan.bins[p];
        uint32_t* status = ws.status + uint64_t{parts} * plan.bins[p];
        const uint32_t tb = hp.tableBits[p];
        if constexpr (MODE == KeyMode::Plain) {
            pd.shift = bitLo + plan.shift[p];
            launchPass<PassKind::Plain64>(tb, parts, stream, kin, vin, kout, vout, n, nDevice, pd, offsets, status,
                                          ws.error);
        } else if (!narrow) {
            launchPass<PassKind::TileDepth64>(tb, parts, stream, kin, vin, kout, vout, n, nDevice, pd, offsets,
                                              status, ws.error);
        } else if (p == 0) {
            launchPass<PassKind::TileNarrow>(tb, parts, stream, kin, vin, kout, vout, n, nDevice, pd, offsets,
                                             status, ws.error);
        } else {
            pd.shift = plan.shift[p] - dropped;
            launchPass<PassKind::Narrow32>(tb, parts, stream, kin, vin, kout, vout, n, nDevice, pd, offsets, status,
                                           ws.error);
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

SortPlan sortPlan(uint32_t keyBits, uint32_t radixBits) {
    SortPlan plan;
    keyBits = std::min(keyBits, 64u);
    if (keyBits == 0) return plan;
    uint32_t radix = radixBits;
    if (radix == 0) radix = (keyBits + 8) / 9 < (keyBits + 7) / 8 ? 9u : 8u;
    radix = std::clamp(radix, kMinTableBits, kMaxRadixBits);
    plan.passes = (keyBits + radix - 1) / radix;
    const uint32_t width = keyBits / plan.passes, wider = keyBits % plan.passes;
    uint32_t shift = 0, bins = 0;
    for (uint32_t p = 0; p < plan.passes; ++p) {
        plan.width[p] = width + (p < wider ? 1u : 0u);
        plan.shift[p] = shift;
        plan.bins[p] = bins;
        shift += plan.width[p];
        bins += 1u << std::max(plan.width[p], kMinTableBits);
    }
    plan.totalBins = bins;
    return plan;
}

uint32_t sortPasses(uint32_t keyBits, uint32_t radixBits) { return sortPlan(keyBits, radixBits).passes; }

bool sortNarrows(KeyMode mode, uint32_t keyBits, uint32_t radixBits) {
    return mode == KeyMode::TileDepth && keyBits > 0 && keyBits <= 32 + sortPlan(keyBits, radixBits).width[0];
}

uint32_t sortPartitions(uint32_t n) { return (n + kPart - 1) / kPart; }

uint64_t sortWorkspaceWords(uint32_t n, uint32_t keyBits, uint32_t radixBits) {
    const uint64_t bins = sortPlan(keyBits, radixBits).totalBins;
    return uint64_t{histogramRows(n)} * bins + bins + uint64_t{sortPartitions(n)} * bins;
}

uint64_t sortWorkspaceWordsUpTo(uint32_t n, uint32_t maxKeyBits, uint32_t radixBits) {
    uint64_t most = 0;
    for (uint32_t b = 1; b <= std::min(maxKeyBits, 64u); ++b) most = std::max(most, sortWorkspaceWords(n, b, radixBits));
    return most;
}

SortWorkspace sortWorkspace(uint32_t* work, uint32_t n, uint32_t keyBits, uint32_t* error, uint32_t radixBits) {
    const uint64_t bins = sortPlan(keyBits, radixBits).totalBins;
    const uint64_t histWords = uint64_t{histogramRows(n)} * bins;
    return SortWorkspace{work, work + histWords, work + histWords + bins, error};
}

bool sortPairs(uint64_t* keys, uint32_t* values, uint64_t* keysAlt, uint32_t* valuesAlt, uint32_t n, KeyMode mode,
               uint32_t bitLo, uint32_t keyBits, const SortWorkspace& ws, hipStream_t stream,
               hipEvent_t afterHistogram, const uint32_t* nDevice, TileDepthLayout depth, uint32_t* alsoZero,
               uint64_t alsoZeroWords, uint32_t radixBits) {
    if ((nDevice == nullptr && n < 2) || n == 0 || keyBits == 0) {
        if (alsoZero != nullptr) fill32(alsoZero, 0u, alsoZeroWords, stream);
        if (afterHistogram != nullptr) (void)hipEventRecord(afterHistogram, stream);
        return false;
    }
    const SortPlan plan = sortPlan(keyBits, radixBits);
    if (mode == KeyMode::TileDepth) {
        launchPasses<KeyMode::TileDepth>(keys, values, keysAlt, valuesAlt, n, nDevice, bitLo, keyBits, plan, depth,
                                         ws, stream, afterHistogram, alsoZero, alsoZeroWords);
    } else {
        launchPasses<KeyMode::Plain>(keys, values, keysAlt, valuesAlt, n, nDevice, bitLo, keyBits, plan, depth, ws,
                                     stream, afterHistogram, alsoZero, alsoZeroWords);
    }
    return (plan.passes & 1u) != 0u;
}

}   // namespace lrt::kern
