// Preprocess: one thread per splat. EWA projection, the optional Mip-Splatting
// 3D filter and the 2D dilation with its opacity compensation, conic, cutoff,
// tile rectangle, the exact count of tiles the footprint touches, and the
// colour from the fp16 harmonics. Follows the reference's splatProject
// (shaders/lrt/splat/splat_project.slang:52-221; relighting, which is off in
// every frame this path draws, is not carried).
//
// Bandwidth: reads 32 B of packed splat (+ 4 B filter) per splat, and
// 4 * shWords B of harmonics (degree > 0) per drawn splat; writes 8 + 4 B per
// splat and 48 B per drawn one. A workgroup first finds which of its 256 splats
// are drawn (the footprint needs no harmonics), then stages only those splats'
// harmonic rows through LDS with coalesced loads, so the per-thread 92-byte
// rows cost whole cache lines once and culled splats cost none.
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "projection.h"
#include "splat_math.h"

namespace lrt::kern {
namespace {

using namespace lrt::dev;

constexpr uint32_t kBlock = kProjectBlock;
constexpr uint32_t kMaxShWords = 23;   // 45 halves at degree 3, rgb per basis

// The coefficient words of the drawn splats, staged: row r of sSh holds the
// words of the workgroup's r-th drawn splat (in splat order).
__device__ __forceinline__ void stageDrawnRows(const uint32_t* __restrict__ sh, uint32_t first, uint32_t words,
                                               uint32_t rows, const uint32_t* sRowOf, uint32_t* sSh) {
    if (words == 0) return;
    const uint32_t total = rows * words;
    // Word w = tid + k * kBlock is word w % words of row w / words; both are
    // stepped, not divided, per k. kBatch loads are issued before any is
    // stored, so their latencies overlap instead of adding up.
    constexpr uint32_t kBatch = 8;
    uint32_t row = threadIdx.x / words, col = threadIdx.x - row * words;
    const uint32_t rowStep = kBlock / words, colStep = kBlock - rowStep * words;
    for (uint32_t w0 = threadIdx.x; w0 < total; w0 += kBlock * kBatch) {
        uint32_t v[kBatch];
#pragma unroll
        for (uint32_t k = 0; k < kBatch; ++k) {
            if (w0 + k * kBlock < total) v[k] = sh[static_cast<uint64_t>(first + sRowOf[row]) * words + col];
            row += rowStep;
            col += colStep;
            if (col >= words) {
                col -= words;
                row += 1;
            }
        }
#pragma unroll
        for (uint32_t k = 0; k < kBatch; ++k) {
            if (w0 + k * kBlock < total) sSh[w0 + k * kBlock] = v[k];
        }
    }
}

// One workgroup's 256 splats of one instance, from the frame constants f.
__device__ __forceinline__ void projectBlock(const FrameParams& f, const float4* __restrict__ positions,
                                             const uint4* __restrict__ shape, const uint32_t* __restrict__ sh,
                                             const float* __restrict__ filter3d, float4* __restrict__ proj,
                                             uint2* __restrict__ rects, uint32_t* __restrict__ touchedOut,
                                             DepthBlocks depth) {
    __shared__ uint32_t sSh[kBlock * kMaxShWords];
    __shared__ uint32_t sRowOf[kBlock];
    __shared__ uint8_t sGraded[kBlock];
    __shared__ uint32_t sDrawn[kBlock / 64];
    __shared__ uint32_t sDepth[2 * (kBlock / 64)];
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u, wave = tid >> 6;
    const uint32_t first = blockIdx.x * kBlock;
    const uint32_t i = first + tid;
    const bool live = i < f.count;
    const uint32_t slot = f.base + i;
    const uint32_t degree = degreeOf(f.restPerColour) < f.shLimit ? degreeOf(f.restPerColour) : f.shLimit;
    const uint32_t words = degree == 0 ? 0u : f.shWords;
    const uint2 emptyRect = make_uint2(1u | (1u << 16), 0u);

    // The footprint: position and shape only.
    uint4 row = make_uint4(0u, 0u, 0u, 0u);
    EditFate fate{true, false, 0.0f};
    SplatShape shapeOf{0.0f, 0.0f, 0.0f, 0.0f, 1.0f, 1.0f, 1.0f, {0.0f, 0.0f, 0.0f, 1.0f}};
    Footprint fp{0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, {1u, 1u, 0u, 0u}};
    bool candidate = false;
    if (live) {
        const float4 p = positions[i];
        row = shape[i];
        // Issued with the two loads above: one trip to memory, not two.
        const float filterOf = f.useFilter3d ? filter3d[i] : 0.0f;
        const float sx = expF(lowHalf(row.y)), sy = expF(highHalf(row.y)), sz = expF(lowHalf(row.z));
        fate = editFate(f.edit, p.x, p.y, p.z, p.w, sx, sy, sz);
        if (!fate.removed) {
            shapeOf = SplatShape{p.x, p.y, p.z, fate.opacity, sx, sy, sz, {0.0f, 0.0f, 0.0f, 0.0f}};
            if (f.useFilter3d) applyFilter3d(filterOf, shapeOf.sx, shapeOf.sy, shapeOf.sz, shapeOf.opacity);
            decodeQuaternion(row.x, shapeOf.q);
            candidate = footprintOf(f, shapeOf, fp);
        }
    }
    // Every lane of the wave counts together, each rectangle among the wave's.
    const uint32_t touched = countTouchedTilesWave(f.width, f.height, fp.cx, fp.cy, fp.ca, fp.cb, fp.cc, fp.cutoff,
                                                   fp.rect, candidate, f.tileShift);

    // The drawn splats in splat order: row `at` of the staged harmonics.
    const bool drawn = touched > 0;
    const uint64_t drawnLanes = __ballot(drawn);
    if (lane == 0) sDrawn[wave] = static_cast<uint32_t>(__popcll(drawnLanes));
    if (depth.entries != nullptr) waveDepthRange(drawn, fp.tz, sDepth);
    __syncthreads();
    storeDepthRange(depth, sDepth);
    uint32_t at = static_cast<uint32_t>(__popcll(drawnLanes & ((1ull << lane) - 1ull)));
    uint32_t rows = 0;
#pragma unroll
    for (uint32_t w = 0; w < kBlock / 64; ++w) {
        if (w < wave) at += sDrawn[w];
        rows += sDrawn[w];
    }
    if (drawn) {
        sRowOf[at] = tid;
        sGraded[at] = fate.graded ? 1u : 0u;
    }
    __syncthreads();
    stageDrawnRows(sh + static_cast<uint64_t>(first) * words, 0u, words, rows, sRowOf, sSh);
    __syncthreads();

    // The colours: thread r takes the workgroup's r-th drawn splat (harmonic
    // row r), so the harmonics run on full waves however many splats were
    // culled, and writes the record's colour; the splat's own thread writes
    // the rest of the record.
    if (tid < rows) {
        const uint32_t j = first + sRowOf[tid];
        const float4 p = positions[j];
        const uint4 own = shape[j];
        // Unpacked with constant indices so the 45 coefficients stay in registers.
        const uint32_t* shRow = sSh + tid * words;
        float coeff[45];
        const uint32_t halves = degree == 0 ? 0u : (degree == 1 ? 9u : degree == 2 ? 24u : 45u);
#pragma unroll
        for (uint32_t k = 0; k < 45; ++k) {
            const uint32_t wd = k < halves ? shRow[k >> 1] : 0u;
            coeff[k] = (k & 1u) == 0u ? lowHalf(wd) : highHalf(wd);
        }
        const SplatShape centre{p.x, p.y, p.z, 0.0f, 0.0f, 0.0f, 0.0f, {0.0f, 0.0f, 0.0f, 1.0f}};
        float dx, dy, dz;
        viewDirection(f, centre, dx, dy, dz);
        float rest[3];
        evaluateRest(coeff, degree, dx, dy, dz, rest);
        float colour[3];
        colour[0] = fmaxS(highHalf(own.z) + rest[0], 0.0f);
        colour[1] = fmaxS(lowHalf(own.w) + rest[1], 0.0f);
        colour[2] = fmaxS(highHalf(own.w) + rest[2], 0.0f);
        if (sGraded[tid] != 0) editGrade(f.edit, colour);
        proj[static_cast<uint64_t>(f.base + j) * 3 + 2] = make_float4(colour[0], colour[1], colour[2], 0.0f);
    }

    if (!live) return;
    if (!drawn) {
        touchedOut[slot] = 0;
        rects[slot] = emptyRect;
        return;
    }
    writeFootprint(fp, proj + static_cast<uint64_t>(slot) * 3, rects + slot);
    touchedOut[slot] = touched;
}

__global__ __launch_bounds__(kBlock) void projectKernel(FrameParams f, const float4* __restrict__ positions,
                                                        const uint4* __restrict__ shape,
                                                        const uint32_t* __restrict__ sh,
                                                        const float* __restrict__ filter3d,
                                                        float4* __restrict__ proj, uint2* __restrict__ rects,
                                                        uint32_t* __restrict__ touchedOut, DepthBlocks depth,
                                                        ZeroSpans zero) {
    // The side clears: grid-stride, so any grid covers them.
    for (uint32_t w = blockIdx.x * kBlock + threadIdx.x; w < zero.aWords || w < zero.bWords; w += gridDim.x * kBlock) {
        if (w < zero.aWords) zero.a[w] = 0u;
        if (w < zero.bWords) zero.b[w] = 0u;
    }
    projectBlock(f, positions, shape, sh, filter3d, proj, rects, touchedOut, depth);
}

constexpr uint32_t kViewsStaged = 32;   // view constants held in LDS at a time
static_assert(sizeof(FrameParams) % 4 == 0, "FrameParams is staged as words");

// One instance seen by many views, splat-major: each thread decodes its splat
// once (scales, edit, 3D filter, rotation, the 45 harmonic coefficients) and
// projects it for every view, the views' constants staged in LDS. The splat's
// bytes are read once for the whole batch instead of once a view.
__global__ __launch_bounds__(kBlock) void projectViewsKernel(const FrameParams* __restrict__ params, uint32_t views,
                                                             const float4* __restrict__ positions,
                                                             const uint4* __restrict__ shape,
                                                             const uint32_t* __restrict__ sh,
                                                             const float* __restrict__ filter3d,
                                                             float4* __restrict__ proj, uint2* __restrict__ rects,
                                                             uint32_t* __restrict__ touchedOut) {
    constexpr uint32_t kWords = sizeof(FrameParams) / 4;
    __shared__ uint32_t sSh[kBlock * kMaxShWords];
    __shared__ uint32_t sFWords[kViewsStaged * kWords];   // FrameParams as words (no constructor in LDS)
    const FrameParams* sF = reinterpret_cast<const FrameParams*>(sFWords);
    const FrameParams& f0 = params[0];   // what every view shares: count, harmonics, edit, filter
    const uint32_t tid = threadIdx.x;
    const uint32_t first = blockIdx.x * kBlock;
    const uint32_t i = first + tid;
    const uint32_t count = f0.count;
    const uint32_t degree = degreeOf(f0.restPerColour) < f0.shLimit ? degreeOf(f0.restPerColour) : f0.shLimit;
    const uint32_t words = degree == 0 ? 0u : f0.shWords;
    if (words > 0) {
        const uint32_t inBlock = count - first < kBlock ? count - first : kBlock;
        const uint32_t total = inBlock * words;
        const uint32_t* src = sh + static_cast<uint64_t>(first) * words;
        constexpr uint32_t kBatch = 8;   // loads in flight before the first store
        for (uint32_t w0 = tid; w0 < total; w0 += kBlock * kBatch) {
            uint32_t v[kBatch];
#pragma unroll
            for (uint32_t b = 0; b < kBatch; ++b) {
                if (w0 + b * kBlock < total) v[b] = src[w0 + b * kBlock];
            }
#pragma unroll
            for (uint32_t b = 0; b < kBatch; ++b) {
                if (w0 + b * kBlock < total) sSh[w0 + b * kBlock] = v[b];
            }
        }
    }
    const bool live = i < count;
    const uint2 emptyRect = make_uint2(1u | (1u << 16), 0u);
    float4 p = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    uint4 row = make_uint4(0u, 0u, 0u, 0u);
    EditFate fate{true, false, 0.0f};
    SplatShape shapeOf{0.0f, 0.0f, 0.0f, 0.0f, 1.0f, 1.0f, 1.0f, {0.0f, 0.0f, 0.0f, 1.0f}};
    if (live) {
        p = positions[i];
        row = shape[i];
        const float filterOf = f0.useFilter3d ? filter3d[i] : 0.0f;   // issued with the loads above
        const float sx = expF(lowHalf(row.y)), sy = expF(highHalf(row.y)), sz = expF(lowHalf(row.z));
        fate = editFate(f0.edit, p.x, p.y, p.z, p.w, sx, sy, sz);
        shapeOf = SplatShape{p.x, p.y, p.z, fate.opacity, sx, sy, sz, {0.0f, 0.0f, 0.0f, 0.0f}};
        if (f0.useFilter3d) applyFilter3d(filterOf, shapeOf.sx, shapeOf.sy, shapeOf.sz, shapeOf.opacity);
        decodeQuaternion(row.x, shapeOf.q);
    }
    __syncthreads();
    float coeff[45];
    {
        const uint32_t halves = degree == 0 ? 0u : (degree == 1 ? 9u : degree == 2 ? 24u : 45u);
        const uint32_t* shRow = sSh + tid * words;
#pragma unroll
        for (uint32_t k = 0; k < 45; ++k) {
            const uint32_t wd = live && k < halves ? shRow[k >> 1] : 0u;
            coeff[k] = (k & 1u) == 0u ? lowHalf(wd) : highHalf(wd);
        }
    }
    const auto colourOf = [&](float dx, float dy, float dz, float colour[3]) {
        float rest[3];
        evaluateRest(coeff, degree, dx, dy, dz, rest);
        colour[0] = fmaxS(highHalf(row.z) + rest[0], 0.0f);
        colour[1] = fmaxS(lowHalf(row.w) + rest[1], 0.0f);
        colour[2] = fmaxS(highHalf(row.w) + rest[2], 0.0f);
        if (fate.graded) editGrade(f0.edit, colour);
    };
    for (uint32_t v0 = 0; v0 < views; v0 += kViewsStaged) {
        const uint32_t n = views - v0 < kViewsStaged ? views - v0 : kViewsStaged;
        __syncthreads();   // the last group's constants are done with
        const uint32_t* src = reinterpret_cast<const uint32_t*>(params + v0);
        for (uint32_t w = tid; w < n * kWords; w += kBlock) sFWords[w] = src[w];
        __syncthreads();
        for (uint32_t v = 0; v < n; ++v) {   // every lane stays for the wave's tile count
            const FrameParams& f = sF[v];
            const uint32_t slot = f.base + i;
            uint2 rect = emptyRect;
            float tz;
            const uint32_t touched = projectWave(f, shapeOf, live && !fate.removed, colourOf,
                                                 proj + static_cast<uint64_t>(slot) * 3, &rect, &tz);
            if (!live) continue;
            rects[slot] = touched > 0 ? rect : emptyRect;
            touchedOut[slot] = touched;
        }
    }
}

}   // namespace

void project(const FrameParams& f, const float4* positions, const uint4* shape, const uint32_t* sh,
             const float* filter3d, float4* proj, uint2* rects, uint32_t* touched, DepthBlocks depth,
             hipStream_t stream, ZeroSpans zero) {
    if (f.count == 0) return;
    const uint32_t blocks = projectBlocks(f.count);
    projectKernel<<<blocks, kBlock, 0, stream>>>(f, positions, shape, sh, filter3d, proj, rects, touched, depth,
                                                 zero);
}

void projectViews(const FrameParams* params, uint32_t views, uint32_t count, const float4* positions,
                  const uint4* shape, const uint32_t* sh, const float* filter3d, float4* proj, uint2* rects,
                  uint32_t* touched, hipStream_t stream) {
    if (count == 0 || views == 0) return;
    projectViewsKernel<<<(count + kBlock - 1) / kBlock, kBlock, 0, stream>>>(params, views, positions, shape, sh,
                                                                            filter3d, proj, rects, touched);
}

}   // namespace lrt::kern
