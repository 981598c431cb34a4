// Host-side launchers of the HIP kernels. Each takes device pointers and the
// stream to queue on; none synchronises. Sizes are validated by the caller
// (rasterizer.hip) before a launch: every launcher assumes its buffers hold
// what its comment says.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "frame.h"

namespace lrt::kern {

// u32 fill of `count` words (used for counters, flags and the look-back
// status words, instead of hipMemsetAsync, so that a captured graph replays it).
void fill32(uint32_t* dst, uint32_t value, uint64_t count, hipStream_t stream);

// ---- scan ----------------------------------------------------------------
// Exclusive prefix sum of `n` u32 into `out` (may not alias `in`), the total
// into *total (saturating at 2^32 - 1) and the number of non-zero inputs into
// *nonzero. `status` holds scanStatusWords(n) u64 words, cleared here. With
// depthRange non-null, the depth-key entries [0, depthEntries) of the frame's
// projecting workgroups (DepthBlocks) are reduced into depthRange[0] (least)
// and depthRange[1] (greatest). alsoZero's first alsoZeroWords words are
// zeroed by the same launch that clears the look-back words.
uint32_t scanPartitions(uint32_t n);
inline uint64_t scanStatusWords(uint32_t n) { return uint64_t{scanPartitions(n)} * 3 / 2 + 2; }
// With zeroed set, the caller has already cleared the look-back words (the
// first scanZeroWords(n) u32 words of status) and alsoZero with a launch that
// runs anyway, and the scan launches no clear of its own.
void scanExclusive(const uint32_t* in, uint32_t* out, uint32_t n, uint64_t* status, uint32_t* total,
                   uint32_t* nonzero, uint32_t* error, hipStream_t stream, const uint32_t* depthBlocks = nullptr,
                   uint32_t depthEntries = 0, uint32_t* depthRange = nullptr, uint32_t* alsoZero = nullptr,
                   uint32_t alsoZeroWords = 0, bool zeroed = false);
// The u32 words of `status` a scan of n elements needs zeroed before it runs.
inline uint32_t scanZeroWords(uint32_t n) { return scanPartitions(n) * 2; }

// Words a launch zeroes on the side (either span may be empty): a later
// launch's look-back words and counters, cleared by one that runs anyway.
struct ZeroSpans {
    uint32_t* a = nullptr;
    uint32_t aWords = 0;
    uint32_t* b = nullptr;
    uint32_t bWords = 0;
};
// The same scan with its scratch in a static device array instead of `status`
// and counters: no allocation on any call. n at most kStaticScanMax
// (hipErrorInvalidValue past it). Calls that overlap on one device share that
// array and race.
constexpr uint32_t kStaticScanMax = 1u << 28;
hipError_t scanExclusiveStatic(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* total, hipStream_t stream);

// ---- projection -----------------------------------------------------------
// One instance: f.count splats from index 0 of the cloud arrays, written to
// slots f.base + i of proj (3 float4 a slot), rects (uint2 a slot: x0 | y0 << 16,
// x1 | y1 << 16) and touched (u32 a slot); each of its projectBlocks(f.count)
// workgroups' depth-key range into `depth`.
void project(const FrameParams& f, const float4* positions, const uint4* shape, const uint32_t* sh,
             const float* filter3d, float4* proj, uint2* rects, uint32_t* touched, DepthBlocks depth,
             hipStream_t stream, ZeroSpans zero = {});
// The same instance (count splats) for `views` views at once: view v's
// constants are params[v] (device), its slots start at params[v].base.
void projectViews(const FrameParams* params, uint32_t views, uint32_t count, const float4* positions,
                  const uint4* shape, const uint32_t* sh, const float* filter3d, float4* proj, uint2* rects,
                  uint32_t* touched, hipStream_t stream);
// One point instance as opaque discs: f.count points (positions float4 xyz 1;
// colours two words a point, f16 r | g, b | a, 8-byte aligned, read only
// without style.constantColour) into slots f.base + i, in the same record,
// rectangle and count layout as a splat's, colour.w = 1; depth-key ranges into
// `depth` as project's are.
void projectPoints(const FrameParams& f, const PointParams& style, const float4* positions, const uint32_t* colours,
                   float4* proj, uint2* rects, uint32_t* touched, DepthBlocks depth, hipStream_t stream);

// ---- duplication ------------------------------------------------------------
// For each of `slots` slots with touched > 0: its pairs at offsets[slot].., keys
// (tile << 32 | bits(view z)) and values (slot). With pairSlot non-null the
// value is instead the pair's own index and pairSlot[index] its slot. Writes
// past `capacity` are dropped and counted in *overflow. With depthRange
// non-null, a slot whose depth key lies outside [depthRange[0], depthRange[1]]
// is counted in *outOfRange (the sort's key layout would not hold it).
void duplicate(const FrameParams& f, uint32_t slots, const float4* proj, const uint2* rects,
               const uint32_t* touched, const uint32_t* offsets, uint64_t* keys, uint32_t* values,
               uint32_t* pairSlot, uint32_t capacity, uint32_t* overflow, hipStream_t stream,
               const uint32_t* depthRange = nullptr, uint32_t* outOfRange = nullptr);
// A batch of views (device ViewInfo array): slot s is view s / splatsPerView's,
// its keys carry the view's tileBase; pairSlot as duplicate()'s. With
// outOfRange non-null, a drawn slot whose depth key lies outside its view's
// [depthLo, depthHi] is counted there.
void duplicateViews(const ViewInfo* views, uint32_t splatsPerView, uint32_t tileShift, uint32_t slots,
                    const float4* proj,
                    const uint2* rects, const uint32_t* touched, const uint32_t* offsets, uint64_t* keys,
                    uint32_t* values, uint32_t* pairSlot, uint32_t capacity, uint32_t* overflow,
                    uint32_t* outOfRange, hipStream_t stream);

// ---- sort -------------------------------------------------------------------
// Which bits of a 64-bit key the digits come from.
enum class KeyMode : uint32_t {
    Plain = 0,        // bits [bitLo, bitLo + keyBits)
    TileDepth = 1,    // (key >> 32) << depthBits | ((key & 0x7FFFFFFF) >> 7) - depthMin:
                      // tile id over the top 24 bits of a positive float depth, less
                      // the frame's least; keyBits = tileBits + depthBits
};

// TileDepth's depth field: the frame's depth keys lie in [depthMin, depthMin +
// 2^depthBits), so the virtual key orders as (tile, depth key) does in fewer
// bits. The default is the whole 24-bit field.
struct TileDepthLayout {
    uint32_t depthMin = 0;
    uint32_t depthBits = 24;
};

// A sort's digits. radixBits (8, 9 or 10; 0 picks 9 where that saves a pass
// and 8 otherwise) bounds a digit: the key takes ceil(keyBits / radixBits)
// passes of keyBits / passes bits, the first keyBits % passes of them one bit
// wider (33 bits: 9 + 8 + 8 + 8 at radix 9, 7 + 7 + 7 + 6 + 6 at radix 8).
// Pass p's digit is bits [shift[p], shift[p] + width[p]) of the key; its digit
// tables hold 2^max(8, width[p]) entries, at bins[p] of a row of totalBins.
constexpr uint32_t kMaxSortPasses = 8;
struct SortPlan {
    uint32_t passes = 0, totalBins = 0;
    uint32_t width[kMaxSortPasses] = {}, shift[kMaxSortPasses] = {}, bins[kMaxSortPasses] = {};
};
SortPlan sortPlan(uint32_t keyBits, uint32_t radixBits);
uint32_t sortPasses(uint32_t keyBits, uint32_t radixBits = 0);

// A TileDepth sort whose key exceeds 32 bits by no more than its first digit
// leaves its keys as u32 (in the u64 buffers' first half-words): the virtual
// key shifted right by narrowShift(keyBits), the bits that first pass consumes,
// so the tile id is the stored key >> (depthBits - narrowShift(keyBits)).
// Wider TileDepth sorts and Plain sorts keep u64 keys.
bool sortNarrows(KeyMode mode, uint32_t keyBits, uint32_t radixBits = 0);
constexpr uint32_t narrowShift(uint32_t keyBits) { return keyBits > 32 ? keyBits - 32 : 0u; }

struct SortWorkspace {
    uint32_t* histogram;   // one row of the plan's totalBins digit counts per histogram workgroup
    uint32_t* offsets;     // totalBins
    uint32_t* status;      // partitions * totalBins
    uint32_t* error;       // 1
};
uint32_t sortPartitions(uint32_t n);
// Workspace words for a sort of capacity n on keyBits bits (up to maxKeyBits
// bits for the second form: the most any key width up to it needs), and its
// layout in `work` (which holds at least that many words).
uint64_t sortWorkspaceWords(uint32_t n, uint32_t keyBits, uint32_t radixBits = 0);
uint64_t sortWorkspaceWordsUpTo(uint32_t n, uint32_t maxKeyBits, uint32_t radixBits = 0);
SortWorkspace sortWorkspace(uint32_t* work, uint32_t n, uint32_t keyBits, uint32_t* error, uint32_t radixBits = 0);
// Stable LSD radix sort, one onesweep launch a pass (digits as sortPlan lays
// them out). The result is in (keys, values) after an even number of passes
// and in (keysAlt, valuesAlt) after an odd one; returns true for the latter.
// The result's keys are u32 when sortNarrows(mode, keyBits, radixBits).
// afterHistogram (may be null) is recorded between the histogram and the passes.
// alsoZero's first alsoZeroWords words are zeroed before the passes (by the
// histogram launch, or by a fill when there is nothing to sort).
// With nDevice non-null the count is *nDevice, read by the kernels, and n is
// the capacity that bounds it (the launch is sized for n; nothing waits on the host).
bool sortPairs(uint64_t* keys, uint32_t* values, uint64_t* keysAlt, uint32_t* valuesAlt, uint32_t n,
               KeyMode mode, uint32_t bitLo, uint32_t keyBits, const SortWorkspace& ws, hipStream_t stream,
               hipEvent_t afterHistogram = nullptr, const uint32_t* nDevice = nullptr,
               TileDepthLayout depth = {}, uint32_t* alsoZero = nullptr, uint64_t alsoZeroWords = 0,
               uint32_t radixBits = 0);

// ---- load: raw splat records into packed splats --------------------------------
// The decode's view of an io::SplatEncoding: field offsets in floats, modes
// numbered as the encoding numbers them, the rest basis functions kept and the
// packed row's words.
struct DecodeParams {
    uint32_t count, stride, keep, shWords;
    uint32_t x, y, z, opacity, scale0, scale1, scale2, rotW, rotX, rotY, rotZ, dc0, dc1, dc2;
    uint32_t restBase, filePerColour, restColourOuter;
    uint32_t opacityMode, scaleMode, colourMode, rotationMode, restMode;
    float positionScale;   // positions are multiplied by it (SPZ: 2^-fractional bits)
    uint32_t flipYZ;       // the file is right-up-back: y and z change sign
};
// Validates records [0, p.count) of `raw` (device), scans the flags into
// `dest` (kept count into *kept) and packs each kept record to base + its
// place. valid/dest hold p.count u32, scanStatus scanStatusWords(p.count) u64.
void decodeSplats(const float* raw, const DecodeParams& p, uint32_t* valid, uint32_t* dest, uint64_t* scanStatus,
                  uint32_t* kept, uint32_t* nonzero, uint32_t* scanError, uint32_t base, float4* positions,
                  uint4* shape, uint32_t* sh, hipStream_t stream);
// box[0..2] = min x, y, z and box[3..5] = max over positions [0, count);
// partial holds boundsPartials(count) * 6 floats.
uint32_t boundsPartials(uint32_t count);
void splatBounds(const float4* positions, uint32_t count, float* partial, float* box, hipStream_t stream);

// ---- ranges -----------------------------------------------------------------
// ranges[t] = [start, end) of tile t's run in the sorted keys, (0, 0) when empty.
// With nDevice non-null the count is *nDevice and `pairs` its capacity.
// zeroed: the caller has already set every range to (0, 0).
void tileRanges(const uint64_t* keys, uint32_t pairs, uint32_t tiles, uint2* ranges, hipStream_t stream,
                const uint32_t* nDevice = nullptr, bool zeroed = false);
// The same over a narrowed sort's u32 keys: the tile is key >> tileShift.
void tileRangesNarrow(const uint32_t* keys, uint32_t tileShift, uint32_t pairs, uint32_t tiles, uint2* ranges,
                      hipStream_t stream, const uint32_t* nDevice = nullptr, bool zeroed = false);
// u64 keys tile << 32 | bits(view z) rebuilt from a narrowed sort's u32 keys
// and the pairs' slots (the view z read from each slot's record), for readback.
void widenKeys(const uint32_t* keys, uint32_t tileShift, const uint32_t* slots, const float4* proj, uint32_t pairs,
               uint64_t* out, hipStream_t stream);

// ---- blend ------------------------------------------------------------------
// Tile indices, heaviest tile (most pairs) first, for the blend's launch order.
void tileOrder(const uint2* ranges, uint32_t tiles, uint32_t* order, hipStream_t stream);
// Per pixel, front to back over its tile's pairs; one wave per 8x8 quadrant,
// whose LDS list holds `list` (64, 128 or 256) records; tiles in `order`
// (null: index order); slots bounds the pair slots. pairSlots holds the
// sorted pairs' slots, or with slotOfPair non-null their pair indices, each
// pair's slot then slotOfPair[index]. lastCount/finalT/accum may be null;
// accum receives each pixel's raw rgb and depth sum before the linearisation
// and the background.
void blend(const FrameParams& f, uint32_t list, uint32_t slots, const uint32_t* order, const uint2* ranges,
           const uint32_t* pairSlots, const float4* proj, float4* colour, float* depth, uint32_t* lastCount,
           float* finalT, float4* accum, hipStream_t stream, uint32_t* walkSteps = nullptr,
           const uint32_t* slotOfPair = nullptr);
// (walkSteps, when not null, receives the sum over quadrant waves of the list
// records each walked: the (pixel, record) steps are 64 times that.)
// The composite walk: a record with colour.w = 1 (a disc) that reaches a pixel
// ends its walk with its colour behind the splats; with underDepth non-null a
// pixel whose under depth is above 0 ends its walk at the first record at or
// behind that depth, and the under-layer's colour (float4 a pixel) lies behind
// what it blended. Writes colour and depth (may be null) only.
void blendComposite(const FrameParams& f, uint32_t list, uint32_t slots, const uint32_t* order, const uint2* ranges,
                    const uint32_t* pairSlots, const float4* proj, const float4* underColour,
                    const float* underDepth, float4* colour, float* depth, hipStream_t stream);

// A frame whose heavy tiles are split: the tiles heaviest first, each tile of
// more than segmentPairs pairs cut into near-equal segments whose quadrants
// are walked by their own waves (partial walks in partialA/partialB, one
// entry a pixel of a tile in each slot), then composited front to back by one wave a quadrant, a
// pixel that would stop inside a segment walking that segment again, so it
// stops where the one walk would. order holds the tiles; tasks/combine/counts
// the plan, sized for taskCapacity >= tiles + pairs / segmentPairs tasks and
// combineCapacity >= the split tiles (slots: pairs / segmentPairs + split
// tiles). Not for composite or training frames.
struct BlendPlan {
    uint32_t* order;
    uint4* tasks;
    uint4* combine;
    uint32_t* counts;   // 2: tasks, split tiles
    float4* partialA;
    float2* partialB;
    uint32_t taskCapacity, combineCapacity;
};
void blendSplit(const FrameParams& f, uint32_t list, uint32_t segmentPairs, uint32_t slots, const uint2* ranges,
                const uint32_t* pairSlots, const float4* proj, const BlendPlan& plan, float4* colour, float* depth,
                uint32_t* lastCount, float* finalT, hipStream_t stream);

// A batch of views in one launch (device FrameParams and ViewInfo arrays;
// views sorted by tileBase, `tiles` in all): as a grid of quadrant tasks, or
// persistent (at most gridCap workgroups taking tasks from *queue, which is
// cleared here). Every view's tiles are 2^tileShift pixels a side.
// lastCount/finalT/accum (each may be null) hold the batch's pixels
// (ViewInfo.pixelBase); accum as blend()'s.
void blendViews(uint32_t depthMode, uint32_t list, bool persistent, uint32_t gridCap, const FrameParams* params,
                const ViewInfo* views, uint32_t viewCount, uint32_t tiles, uint32_t tileShift, uint32_t slots,
                const uint2* ranges,
                const uint32_t* pairSlots, const float4* proj, uint32_t* lastCount, float* finalT, float4* accum,
                uint32_t* queue, hipStream_t stream, const uint32_t* slotOfPair = nullptr);

// ---- brute force in FP64 -----------------------------------------------------------
// Bytes of one FP64 brute-force record.
uint64_t referenceRecordBytes();
// Instance f's splats in FP64 into records[slot]; keys[slot] = the depth order
// key (exact z bits, or with keyed the rasteriser's 24-bit key in the high
// word), all ones for a culled splat; values[slot] = slot; *live += visible.
void referenceProject(const FrameParams& f, const float4* positions, const uint4* shape, const uint32_t* sh,
                      const float* filter3d, void* records, uint64_t* keys, uint32_t* values, uint32_t* live,
                      bool keyed, hipStream_t stream);
// Every pixel over the first *live slots of `order`, in FP64: colour 4 doubles
// a pixel, depth one (may be null). With pixelList, only the listCount pixels
// it names (y * width + x), output k for pixelList[k].
void referenceBlend(const FrameParams& f, const void* records, const uint32_t* order, const uint32_t* live,
                    double* colour, double* depth, hipStream_t stream, const uint32_t* pixelList = nullptr,
                    uint32_t listCount = 0);

// ---- differentiable path -------------------------------------------------------
// An FP32-parameter cloud (device pointers): means N x 3, log scales N x 3,
// quaternions N x 4 (x, y, z, w; any length), opacity logits N, harmonics
// N x (1 + restPerColour) x 3 (dc first, rgb per basis), filter3d N or null.
struct DiffCloud {
    const float* means;
    const float* logScales;
    const float* quats;
    const float* logits;
    const float* sh;
    const float* filter3d;
    uint32_t count;
    uint32_t restPerColour;
};
// Gradients of the same shapes, overwritten.
struct DiffGrads {
    float* means;
    float* logScales;
    float* quats;
    float* logits;
    float* sh;
};
void projectDiff(const FrameParams& f, const DiffCloud& cloud, float4* proj, uint2* rects, uint32_t* touched,
                 DepthBlocks depth, hipStream_t stream);
// The same cloud for `views` views at once, splat-major: view v's constants
// are params[v] (device), its slots start at params[v].base.
void projectDiffViews(const FrameParams* params, uint32_t views, const DiffCloud& cloud, float4* proj, uint2* rects,
                      uint32_t* touched, hipStream_t stream);
// sortedSlots[i] = pairSlot[pairIndex[i]] for i < pairs (or < *nDevice, pairs
// then being the capacity): the debug readback's sorted slots of a training
// frame (its blends read pairSlot[pairIndex[i]] as they go).
void gatherSlots(const uint32_t* pairIndex, const uint32_t* pairSlot, uint32_t pairs, uint32_t* sortedSlots,
                 hipStream_t stream, const uint32_t* nDevice = nullptr);
// Each pair's nine screen-space gradient terms, 36 bytes a pair, in one buffer
// of `capacity` pairs: a 32-byte record ab[2i], ab[2i + 1] = d (centre x,
// centre y, conic a, conic b), d (conic c, alpha, r, g), then an array c[i] =
// d b. The blend backward writes pairs at scattered places; a whole aligned
// 32-byte record is one full DRAM sector, where two separate float4 arrays
// would leave twice the partly written sectors.
struct PairGrads {
    float4* ab;
    float* c;
};
constexpr uint64_t kPairGradBytes = 36;
inline PairGrads pairGradsIn(void* buffer, uint32_t capacity) {
    auto* ab = static_cast<float4*>(buffer);
    return PairGrads{ab, reinterpret_cast<float*>(ab + 2ull * capacity)};
}
// dL/d(colour) per pixel -> one screen-space gradient record per pair, at the
// pair's index in pairGrad. pairIndex: the sorted pairs' indices; pairSlot:
// each pair's slot, by pair index.
void blendBackward(const FrameParams& f, uint32_t slots, const uint2* ranges, const uint32_t* pairSlot,
                   const uint32_t* pairIndex, const float4* proj, const float4* accum, const uint32_t* lastCount,
                   const float* finalT, const float4* gradColour, PairGrads pairGrad, hipStream_t stream,
                   const uint32_t* order = nullptr);
// The blend backward of every tile of a batch (device FrameParams and ViewInfo
// arrays as blendViews'; accum/lastCount/finalT hold the batch's pixels at
// ViewInfo.pixelBase); gradColours: a device array of one colour-gradient
// pointer a view. order (may be null): the tiles in launch order (tileOrder's).
void blendBackwardViews(const FrameParams* params, const ViewInfo* views, uint32_t viewCount, uint32_t tiles,
                        uint32_t slots, const uint2* ranges, const uint32_t* pairSlot, const uint32_t* pairIndex,
                        const float4* proj, const float4* accum, const uint32_t* lastCount, const float* finalT,
                        const float4* const* gradColours, PairGrads pairGrad, hipStream_t stream,
                        const uint32_t* order = nullptr);
// Each splat's runs of pair records in every view -> gradients of its
// parameters summed over the views in view order.
void projectBackwardViews(const FrameParams* params, uint32_t views, const DiffCloud& cloud, const uint32_t* touched,
                          const uint32_t* offsets, PairGrads pairGrad, const DiffGrads& grads,
                          hipStream_t stream);
// Each splat's run of pair records -> gradients of its parameters.
void projectBackward(const FrameParams& f, const DiffCloud& cloud, const uint32_t* touched, const uint32_t* offsets,
                     PairGrads pairGrad, const DiffGrads& grads, hipStream_t stream);

}   // namespace lrt::kern
