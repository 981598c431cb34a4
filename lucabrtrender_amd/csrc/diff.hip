// The differentiable path: the preprocess from FP32 parameters, and the
// backward of the blend and of the preprocess (EWA projection, scales,
// rotation, opacity, spherical harmonics). The reference has no backward; the
// forward it differentiates is the reference's frame (splat_project.slang:
// 52-221, splat_blend.slang:23-114), drawn by the same kernels as every other
// frame.
//
// Per-pixel state is not stored per splat: the blend leaves each pixel's final
// transmittance, the place of its last contributor in its tile's pair list
// and its raw accumulated rgb. The backward walks each tile's pairs back to
// front from there, recovering T before each contributor as T / (1 - alpha).
//
// Gradients reach the splats without atomics. The duplication gave each pair
// its index in the slot-major pair order as its sort value; a tile's four
// quadrant waves reduce each pair's per-pixel gradients across their lanes
// (DPP), the workgroup adds the four quadrants in LDS and writes one record per
// pair at that index; each splat then sums its own contiguous run of records.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "kernels.h"
#include "projection.h"
#include "splat_math.h"

namespace lrt::kern {
namespace {

using namespace lrt::dev;

constexpr uint32_t kBlock = 256;
constexpr float kShC0 = 0.28209479177387814f;
constexpr float kExpScale = -0.72134752044448170368f;   // -log2(e) / 2

__device__ __forceinline__ uint32_t shDegree(const FrameParams& f) {
    const uint32_t d = degreeOf(f.restPerColour);
    return d < f.shLimit ? d : f.shLimit;
}

// Basis functions used at each degree: 3, 8, 15 rest coefficients per colour.
__device__ __forceinline__ uint32_t restCount(uint32_t degree) {
    return degree == 0 ? 0u : degree == 1 ? 3u : degree == 2 ? 8u : 15u;
}

// A splat of the FP32 cloud, decoded as the packed path decodes its own:
// scales exp(log scale), opacity sigmoid(logit), the quaternion normalised
// (the zero quaternion is the identity).
struct Decoded {
    SplatShape shape;
    float sigmoid;     // opacity before the 3D filter
    float qLen;        // |q| before normalising
    float s[3];        // exp(log scale), before the 3D filter
};

__device__ __forceinline__ Decoded decode(const DiffCloud& c, uint32_t i, bool useFilter3d) {
    Decoded d;
    const float* mean = c.means + static_cast<uint64_t>(i) * 3;
    const float* ls = c.logScales + static_cast<uint64_t>(i) * 3;
    const float* qr = c.quats + static_cast<uint64_t>(i) * 4;
    d.s[0] = expF(ls[0]);
    d.s[1] = expF(ls[1]);
    d.s[2] = expF(ls[2]);
    d.sigmoid = 1.0f / (1.0f + expF(-c.logits[i]));
    float q[4] = {qr[0], qr[1], qr[2], qr[3]};
    d.qLen = __builtin_sqrtf((q[0] * q[0] + q[1] * q[1]) + (q[2] * q[2] + q[3] * q[3]));
    if (d.qLen > 0.0f) {
        for (float& v : q) v = v / d.qLen;
    } else {
        q[0] = q[1] = q[2] = 0.0f;
        q[3] = 1.0f;
    }
    d.shape = SplatShape{mean[0], mean[1], mean[2], d.sigmoid, d.s[0], d.s[1], d.s[2], {q[0], q[1], q[2], q[3]}};
    if (useFilter3d) applyFilter3d(c.filter3d[i], d.shape.sx, d.shape.sy, d.shape.sz, d.shape.opacity);
    return d;
}

// The splat's harmonic coefficients its frames use: dc rgb, and the rest
// bases up to the degree (zero past it).
__device__ __forceinline__ void loadHarmonics(const DiffCloud& c, uint32_t i, uint32_t degree, float dc[3],
                                              float coeff[45]) {
    const float* row = c.sh + static_cast<uint64_t>(i) * ((1u + c.restPerColour) * 3u);
    const uint32_t used = restCount(degree) * 3u;
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) dc[ch] = row[ch];
#pragma unroll
    for (uint32_t k = 0; k < 45; ++k) coeff[k] = k < used ? row[3 + k] : 0.0f;
}

// A workgroup's rows of a row-major float array (stride floats a row, at most
// kRowPitchMax - 1) moved between global memory and LDS, row r at LDS word
// r * pitch with pitch = stride | 1: consecutive threads move consecutive
// words of global memory (coalesced; kBatch words in flight a thread), and a
// thread reading its own row of LDS meets no bank conflict (odd pitch).
constexpr uint32_t kRowPitchMax = 49;

__device__ __forceinline__ uint32_t rowPitch(uint32_t stride) { return stride | 1u; }

template <bool LOAD>
__device__ __forceinline__ void moveRows(std::conditional_t<LOAD, const float*, float*> global, uint32_t stride,
                                         uint32_t rows, float* rowsLds) {
    constexpr uint32_t kBatch = 8;
    const uint32_t total = rows * stride, pitch = rowPitch(stride);
    uint32_t row = threadIdx.x / stride, col = threadIdx.x - row * stride;
    const uint32_t rowStep = kBlock / stride, colStep = kBlock - rowStep * stride;
    for (uint32_t e0 = threadIdx.x; e0 < total; e0 += kBlock * kBatch) {
        uint32_t at[kBatch];
#pragma unroll
        for (uint32_t k = 0; k < kBatch; ++k) {
            at[k] = row * pitch + col;
            row += rowStep;
            col += colStep;
            if (col >= stride) {
                col -= stride;
                row += 1;
            }
        }
        float v[kBatch];
        if constexpr (LOAD) {
#pragma unroll
            for (uint32_t k = 0; k < kBatch; ++k) v[k] = e0 + k * kBlock < total ? global[e0 + k * kBlock] : 0.0f;
#pragma unroll
            for (uint32_t k = 0; k < kBatch; ++k) {
                if (e0 + k * kBlock < total) rowsLds[at[k]] = v[k];
            }
        } else {
#pragma unroll
            for (uint32_t k = 0; k < kBatch; ++k) v[k] = e0 + k * kBlock < total ? rowsLds[at[k]] : 0.0f;
#pragma unroll
            for (uint32_t k = 0; k < kBatch; ++k) {
                if (e0 + k * kBlock < total) global[e0 + k * kBlock] = v[k];
            }
        }
    }
}

static_assert(kBlock == kProjectBlock, "one depth-range entry a workgroup of kProjectBlock slots");

// The footprints first; then the workgroup's harmonic rows, half the
// workgroup's at a time, staged through LDS with coalesced loads (a half with
// no drawn splat loads nothing), and each drawn splat's colour from its row.
__global__ __launch_bounds__(kBlock) void projectDiffKernel(FrameParams f, DiffCloud c, float4* __restrict__ proj,
                                                            uint2* __restrict__ rects,
                                                            uint32_t* __restrict__ touchedOut, DepthBlocks depth) {
    constexpr uint32_t kHalf = kBlock / 2;
    __shared__ float sRows[kHalf * kRowPitchMax];
    __shared__ uint32_t sDepth[2 * (kBlock / 64)];
    const uint32_t blockFirst = blockIdx.x * kBlock;
    const uint32_t i = blockFirst + threadIdx.x;
    const bool live = i < f.count;   // every lane stays for the wave's tile count and the barriers
    const uint32_t slot = f.base + i;
    const Decoded d = decode(c, live ? i : 0u, f.useFilter3d != 0);
    Footprint fp{0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, {1u, 1u, 0u, 0u}};
    const bool candidate = live && footprintOf(f, d.shape, fp);
    const uint32_t touched = countTouchedTilesWave(f.width, f.height, fp.cx, fp.cy, fp.ca, fp.cb, fp.cc, fp.cutoff,
                                                   fp.rect, candidate, f.tileShift);
    if (depth.entries != nullptr) waveDepthRange(touched > 0, fp.tz, sDepth);
    const uint32_t degree = shDegree(f);
    const uint32_t stride = (1u + c.restPerColour) * 3u, pitch = rowPitch(stride);
    const uint32_t rows = f.count - blockFirst < kBlock ? f.count - blockFirst : kBlock;
    const uint32_t half = threadIdx.x / kHalf;
    float colour[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (uint32_t h = 0; h < 2; ++h) {
        // A barrier: the last half's rows are read, and the depth words are written.
        if (!__syncthreads_or(half == h && touched > 0)) continue;
        const uint32_t first = h * kHalf;
        moveRows<true>(c.sh + static_cast<uint64_t>(blockFirst + first) * stride, stride,
                       rows - first < kHalf ? rows - first : kHalf, sRows);
        __syncthreads();
        if (half == h && touched > 0) {
            const float* row = sRows + (threadIdx.x - first) * pitch;
            float dx, dy, dz;
            viewDirection(f, d.shape, dx, dy, dz);
            float rest[3];
            evaluateRest(row + 3, degree, dx, dy, dz, rest);
#pragma unroll
            for (int ch = 0; ch < 3; ++ch) colour[ch] = fmaxS((0.5f + kShC0 * row[ch]) + rest[ch], 0.0f);
        }
    }
    if (depth.entries != nullptr) storeDepthRange(depth, sDepth);
    if (!live) return;
    if (touched > 0) {
        writeProjection(fp, colour, proj + static_cast<uint64_t>(slot) * 3, rects + slot);
    } else {
        rects[slot] = make_uint2(1u | (1u << 16), 0u);
    }
    touchedOut[slot] = touched;
}

constexpr uint32_t kViewsStaged = 32;   // view constants held in LDS at a time
static_assert(sizeof(FrameParams) % 4 == 0, "FrameParams is staged as words");

// One cloud seen by many views, splat-major: each thread decodes its splat and
// reads its 48 harmonic coefficients once, then projects it for every view
// (the views' constants staged in LDS), each view's colour by the same
// arithmetic as projectDiffKernel's.
__global__ __launch_bounds__(kBlock) void projectDiffViewsKernel(const FrameParams* __restrict__ params,
                                                                 uint32_t views, DiffCloud c,
                                                                 float4* __restrict__ proj,
                                                                 uint2* __restrict__ rects,
                                                                 uint32_t* __restrict__ touchedOut) {
    constexpr uint32_t kWords = sizeof(FrameParams) / 4;
    __shared__ uint32_t sFWords[kViewsStaged * kWords];   // FrameParams as words (no constructor in LDS)
    const FrameParams* sF = reinterpret_cast<const FrameParams*>(sFWords);
    const FrameParams& f0 = params[0];   // what every view shares: the filter, the harmonics' degree
    const uint32_t tid = threadIdx.x;
    const uint32_t i = blockIdx.x * kBlock + tid;
    const bool live = i < c.count;   // every lane stays for the wave's tile counts
    const Decoded d = decode(c, live ? i : 0u, f0.useFilter3d != 0);
    const uint32_t degree = shDegree(f0);
    float dc[3], coeff[45];
    loadHarmonics(c, live ? i : 0u, degree, dc, coeff);
    const auto colourOf = [&](float dx, float dy, float dz, float colour[3]) {
        float rest[3];
        evaluateRest(coeff, degree, dx, dy, dz, rest);
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) colour[ch] = fmaxS((0.5f + kShC0 * dc[ch]) + rest[ch], 0.0f);
    };
    const uint2 emptyRect = make_uint2(1u | (1u << 16), 0u);
    for (uint32_t v0 = 0; v0 < views; v0 += kViewsStaged) {
        const uint32_t n = views - v0 < kViewsStaged ? views - v0 : kViewsStaged;
        __syncthreads();   // the last group's constants are done with
        const uint32_t* src = reinterpret_cast<const uint32_t*>(params + v0);
        for (uint32_t w = tid; w < n * kWords; w += kBlock) sFWords[w] = src[w];
        __syncthreads();
        for (uint32_t v = 0; v < n; ++v) {
            const FrameParams& f = sF[v];
            const uint32_t slot = f.base + i;
            uint2 rect = emptyRect;
            float tz;
            const uint32_t touched =
                projectWave(f, d.shape, live, colourOf, proj + static_cast<uint64_t>(slot) * 3, &rect, &tz);
            if (!live) continue;
            rects[slot] = touched > 0 ? rect : emptyRect;
            touchedOut[slot] = touched;
        }
    }
}

__global__ __launch_bounds__(kBlock) void gatherSlotsKernel(const uint32_t* __restrict__ pairIndex,
                                                            const uint32_t* __restrict__ pairSlot, uint32_t cap,
                                                            const uint32_t* __restrict__ nDevice,
                                                            uint32_t* __restrict__ sortedSlots) {
    uint32_t pairs = cap;
    if (nDevice != nullptr && *nDevice < cap) pairs = *nDevice;
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < pairs) sortedSlots[i] = pairSlot[pairIndex[i]];
}

// ---- blend backward ---------------------------------------------------------

// DPP lane moves for the wave sum; lanes outside row_mask read 0.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dppMove(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROW_MASK, 0xF, false));
}

// The same move leaving the rows outside row_mask undefined (one instruction
// fewer than reading 0 there): for steps whose other rows no later step reads.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dppMoveRows(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, ROW_MASK, 0xF, false));
}

// Sum over the 64 lanes, landing in lane 63 (the other lanes hold partial or
// undefined values): pairs, quads, half rows and rows by quad_perm and the row
// mirrors, then row 0 into row 1 and row 2 into row 3 (row_bcast:15), then
// rows 0-1 into rows 2-3 (row_bcast:31). Lane 63 reads only rows the earlier
// steps wrote. Every lane must be active.
__device__ __forceinline__ float rowSum(float v) {
    v += dppMove<0xB1, 0xF>(v);       // quad_perm [1, 0, 3, 2]
    v += dppMove<0x4E, 0xF>(v);       // quad_perm [2, 3, 0, 1]
    v += dppMove<0x141, 0xF>(v);      // row_half_mirror
    v += dppMove<0x140, 0xF>(v);      // row_mirror
    return v;                         // every lane holds its row's sum
}

__device__ __forceinline__ float waveSum(float v) {
    v = rowSum(v);
    v += dppMoveRows<0x142, 0xA>(v);  // row_bcast:15 into rows 1 and 3
    v += dppMoveRows<0x143, 0xC>(v);  // row_bcast:31 into rows 2 and 3
    return v;
}

// Two terms' halves (lanes 0-31 and 32-63) summed at once: permlane32_swap
// leaves a's low half and b's low half in one register and both high halves in
// the other, so one add gives a's half sums in lanes 0-31 and b's in 32-63.
__device__ __forceinline__ float halvesOfTwo(float a, float b) {
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(a), __float_as_uint(b), false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}

// The same for rows: permlane16_swap leaves rows 0 and 2 of a and of b in one
// register and rows 1 and 3 in the other, so one add gives row pairs summed:
// rows 0-3 of the result hold a's rows 0+1, b's rows 0+1, a's rows 2+3 and
// b's rows 2+3.
__device__ __forceinline__ float rowsOfTwo(float a, float b) {
    const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(a), __float_as_uint(b), false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}

// A record's nine terms, each summed over the wave's 64 lanes, stored to
// out[0..8] by the lanes that end with them. Terms 0-7 go in pairs through
// halvesOfTwo, then the pairs' results in pairs through rowsOfTwo, which
// leaves four terms a register, one a row (rows 0-3: terms 0, 2, 1, 3 of the
// four), for the in-row steps; term 8 takes the whole-wave sum. 26
// instructions where nine whole-wave sums take 54. Every lane must be active.
__device__ __forceinline__ void waveSum9(const float v[9], float* out) {
    const float w0 = halvesOfTwo(v[0], v[1]), w1 = halvesOfTwo(v[2], v[3]);
    const float w2 = halvesOfTwo(v[4], v[5]), w3 = halvesOfTwo(v[6], v[7]);
    const float x0 = rowSum(rowsOfTwo(w0, w1)), x1 = rowSum(rowsOfTwo(w2, w3));
    const float s8 = waveSum(v[8]);
    const uint32_t lane = __lane_id();
    if ((lane & 15u) == 0u) {
        const uint32_t row = lane >> 4;
        const uint32_t at = ((row & 1u) << 1) | (row >> 1);
        out[at] = x0;
        out[4 + at] = x1;
    }
    if (lane == 63) out[8] = s8;
}

// sRGB decode and its derivative, for the blend's linearisation
// (reference color.slang:7-10).
__device__ __forceinline__ void srgbToLinearGrad(float x, float& value, float& slope) {
    x = fmaxS(x, 0.0f);
    if (x <= 0.04045f) {
        value = x / 12.92f;
        slope = 1.0f / 12.92f;
    } else {
        const float u = (x + 0.055f) / 1.055f;
        value = expF(2.4f * logF(u));
        slope = 2.4f / 1.055f * (value / u);
    }
}

constexpr uint32_t kWaves = 4;
constexpr uint32_t kBatch = 128;   // pairs staged per step, by the first 128 threads

// One tile of frame f (its pairs [range.x, range.y) of the sorted order): the
// workgroup's four waves take its four quadrants. accum, lastCount, finalT and
// gradColour are the frame's own per-pixel arrays.
__device__ __forceinline__ void blendBackwardTile(const FrameParams& f, uint32_t tile, uint2 range, uint32_t slots,
                                                  const uint32_t* __restrict__ pairSlot,
                                                  const uint32_t* __restrict__ pairIndex,
                                                  const float4* __restrict__ proj,
                                                  const float4* __restrict__ accum,
                                                  const uint32_t* __restrict__ lastCount,
                                                  const float* __restrict__ finalT,
                                                  const float4* __restrict__ gradColour,
                                                  PairGrads pairGrad) {
    __shared__ float4 sGeom[kBatch];     // centre.xy, cutoff, alpha
    __shared__ float4 sConic[kBatch];    // a, 2b, c, (unused)
    __shared__ float4 sColour[kBatch];   // rgb, bits of the pair's place in the tile's range
    __shared__ uint32_t sMask[kBatch];   // quadrants the pair can reach
    __shared__ uint16_t sList[kWaves][kBatch];
    __shared__ uint32_t sCount[kWaves][kWaves];
    __shared__ float4 sPart[kWaves][kBatch][3];
    __shared__ uint32_t sMaxLast;

    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u;
    const uint32_t wave = tid >> 6;
    const uint32_t tileX = tile % f.tilesX;
    const uint32_t tileY = tile / f.tilesX;
    const uint32_t px = tileX * kTile + (wave & 1u) * 8u + (lane & 7u);
    const uint32_t py = tileY * kTile + (wave >> 1) * 8u + (lane >> 3);
    const bool inside = px < f.width && py < f.height;
    const float pxf = static_cast<float>(px) + 0.5f;
    const float pyf = static_cast<float>(py) + 0.5f;
    float boxLo[kWaves][2], boxHi[kWaves][2];
#pragma unroll
    for (uint32_t q = 0; q < kWaves; ++q) {
        const uint32_t x0 = tileX * kTile + (q & 1u) * 8u, y0 = tileY * kTile + (q >> 1) * 8u;
        const uint32_t x1 = x0 + 8u < f.width ? x0 + 8u : f.width;
        const uint32_t y1 = y0 + 8u < f.height ? y0 + 8u : f.height;
        boxLo[q][0] = static_cast<float>(x0) + 0.5f;
        boxLo[q][1] = static_cast<float>(y0) + 0.5f;
        boxHi[q][0] = x1 > x0 ? static_cast<float>(x1 - 1u) + 0.5f : -1e30f;
        boxHi[q][1] = y1 > y0 ? static_cast<float>(y1 - 1u) + 0.5f : -1e30f;
    }

    // The pixel's seeds: dL/d(raw rgb) and dL/d(final T) through the
    // linearisation and the background (splat_blend.slang:89-107).
    uint32_t last = 0;
    float tFinal = 1.0f, dR[3] = {0.0f, 0.0f, 0.0f}, dT = 0.0f;
    if (inside) {
        const uint32_t at = py * f.width + px;
        last = lastCount[at];
        tFinal = finalT[at];
        const float4 go = gradColour[at];
        const float4 raw = accum[at];
        const float coverage = 1.0f - tFinal;
        const float g[3] = {go.x, go.y, go.z};
        const float rgb[3] = {raw.x, raw.y, raw.z};
        dT = go.w * (f.bg[3] - 1.0f);
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) {
            dT += g[ch] * f.bg[ch];
            if (f.linearise && coverage > 1e-6f) {
                const float x = rgb[ch] / coverage;
                float value, slope;
                srgbToLinearGrad(x, value, slope);
                dR[ch] = g[ch] * slope;
                dT -= g[ch] * (value - slope * x);   // d coverage / d T = -1
            } else {
                dR[ch] = g[ch];
            }
        }
    }
    if (tid == 0) sMaxLast = 0;
    __syncthreads();
    if (last > 0) atomicMax(&sMaxLast, last);
    __syncthreads();
    // A last contributor lies inside its tile's range; the bound keeps the walk there.
    const uint32_t maxLast = sMaxLast < range.y - range.x ? sMaxLast : range.y - range.x;
    // Pairs past every pixel's last contributor get zero gradients.
    for (uint32_t pos = range.x + maxLast + tid; pos < range.y; pos += kBlock) {
        const uint32_t at = pairIndex[pos];
        pairGrad.ab[2 * uint64_t{at}] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        pairGrad.ab[2 * uint64_t{at} + 1] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        pairGrad.c[at] = 0.0f;
    }

    // t: T before the walk's current pair; behindDot: sum over the channels of
    // dR times the colour accumulated behind it (the only use of that colour).
    float t = tFinal, behindDot = 0.0f;
    const float dTFinal = dT * tFinal;
    const uint64_t below = (1ull << lane) - 1ull;
    for (uint32_t top = range.x + maxLast; top > range.x;) {
        const uint32_t lo = top - range.x > kBatch ? top - kBatch : range.x;
        const uint32_t n = top - lo;
        // 1. stage entry k = the pair at top - 1 - k, with its quadrant mask
        uint32_t mask = 0;
        if (tid < kBatch) {
            float4 screen = make_float4(0.0f, 0.0f, 0.0f, -1.0f), conic = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            float4 col = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            uint32_t place = 0;
            if (tid < n) {
                const uint32_t pos = top - 1u - tid;
                place = pos - range.x;
                const uint32_t slot = pairSlot[pairIndex[pos]];
                if (slot < slots) {
                    const uint64_t s3 = static_cast<uint64_t>(slot) * 3;
                    screen = proj[s3];
                    conic = proj[s3 + 1];
                    col = proj[s3 + 2];
#pragma unroll
                    for (uint32_t q = 0; q < kWaves; ++q) {
                        if (boxHi[q][0] < 0.0f || boxHi[q][1] < 0.0f) continue;
                        if (quadrantReach(conic.x, conic.y, conic.z, screen.w, boxLo[q][0] - screen.x,
                                             boxLo[q][1] - screen.y, boxHi[q][0] - screen.x,
                                             boxHi[q][1] - screen.y)) {
                            mask |= 1u << q;
                        }
                    }
                }
            }
            sGeom[tid] = make_float4(screen.x, screen.y, screen.w, conic.w);
            sConic[tid] = make_float4(conic.x, conic.y + conic.y, conic.z, 0.0f);
            sColour[tid] = make_float4(col.x, col.y, col.z, __uint_as_float(place));
            sMask[tid] = mask;
        }
        // 2. one list per quadrant, back to front
        uint32_t before[kWaves];
#pragma unroll
        for (uint32_t q = 0; q < kWaves; ++q) {
            const uint64_t vote = __ballot((mask >> q) & 1u);
            before[q] = static_cast<uint32_t>(__popcll(vote & below));
            if (lane == 0) sCount[q][wave] = static_cast<uint32_t>(__popcll(vote));
        }
        __syncthreads();
#pragma unroll
        for (uint32_t q = 0; q < kWaves; ++q) {
            if ((mask >> q) & 1u) {
                uint32_t at = before[q];
#pragma unroll
                for (uint32_t w = 0; w + 1 < kWaves; ++w) at += w < wave ? sCount[q][w] : 0u;
                sList[q][at] = static_cast<uint16_t>(tid);
            }
        }
        const uint32_t count = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(
            static_cast<int>(sCount[wave][0] + sCount[wave][1] + sCount[wave][2] + sCount[wave][3])));
        __syncthreads();
        // 3. this quadrant's pixels, pair by pair, back to front
        for (uint32_t j = 0; j < count; ++j) {
            const uint32_t k = sList[wave][j];
            const float4 g0 = sGeom[k];
            const float4 g1 = sConic[k];
            const float4 g2 = sColour[k];
            const uint32_t place = __float_as_uint(g2.w);
            // The forward's arithmetic, so the same pairs pass the same tests.
            const float dx = pxf - g0.x;
            const float dy = pyf - g0.y;
            const float power = __builtin_fmaf(dx, __builtin_fmaf(g1.x, dx, g1.y * dy), (g1.z * dy) * dy);
            const float gauss = __builtin_amdgcn_exp2f(kExpScale * power);
            const float raw = g0.w * gauss;
            const float alpha = __builtin_fminf(0.99f, raw);
            const bool ok = (place < last) & (power <= g0.z) & (alpha >= 1.0f / 255.0f);
            float v[9] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
            if (ok) {
                const float inv = __builtin_amdgcn_rcpf(1.0f - alpha);
                const float tBefore = t * inv;
                // sum_ch dR (T_before c - behind / (1 - alpha)) - dT tFinal / (1 - alpha)
                const float dotColour = __builtin_fmaf(dR[2], g2.z, __builtin_fmaf(dR[1], g2.y, dR[0] * g2.x));
                const float dAlpha = __builtin_fmaf(tBefore, dotColour, -(behindDot + dTFinal) * inv);
                const float weight = tBefore * alpha;
#pragma unroll
                for (int ch = 0; ch < 3; ++ch) v[6 + ch] = dR[ch] * weight;
                behindDot = __builtin_fmaf(weight, dotColour, behindDot);
                t = tBefore;
                if (raw < 0.99f) {
                    // alpha = opacity * exp(-power / 2). The power's gradient
                    // is summed as moments over the pixels (dPower dx, dPower
                    // dy, dPower dx^2, dPower dx dy, dPower dy^2); the pair's
                    // conic turns them into its centre and conic terms once the
                    // quadrants are added.
                    const float dPower = -0.5f * alpha * dAlpha;
                    v[5] = dAlpha * gauss;
                    v[0] = dPower * dx;
                    v[1] = dPower * dy;
                    v[2] = v[0] * dx;
                    v[3] = v[0] * dy;
                    v[4] = v[1] * dy;
                }
            }
            if (__ballot(ok) != 0) {
                waveSum9(v, reinterpret_cast<float*>(&sPart[wave][k][0]));
            } else if (lane == 63) {
                sPart[wave][k][0] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                sPart[wave][k][1] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                sPart[wave][k][2].x = 0.0f;
            }
        }
        __syncthreads();
        // 4. the four quadrants' sums, one record per pair
        if (tid < n) {
            const uint32_t m = sMask[tid];
            float4 acc[2] = {make_float4(0.0f, 0.0f, 0.0f, 0.0f), make_float4(0.0f, 0.0f, 0.0f, 0.0f)};
            float acc8 = 0.0f;
#pragma unroll
            for (uint32_t q = 0; q < kWaves; ++q) {
                if ((m >> q) & 1u) {
#pragma unroll
                    for (int r = 0; r < 2; ++r) {
                        const float4 p = sPart[q][tid][r];
                        acc[r].x += p.x;
                        acc[r].y += p.y;
                        acc[r].z += p.z;
                        acc[r].w += p.w;
                    }
                    acc8 += sPart[q][tid][2].x;
                }
            }
            // The moments to d(centre) = -(2a Sx + 2b Sy, 2b Sx + 2c Sy) and
            // d(conic) = (Sxx, 2 Sxy, Syy); sConic holds (a, 2b, c).
            const float4 cn = sConic[tid];
            const float sx = acc[0].x, sy = acc[0].y;
            acc[0].x = -(2.0f * cn.x * sx + cn.y * sy);
            acc[0].y = -(cn.y * sx + 2.0f * cn.z * sy);
            acc[0].w = 2.0f * acc[0].w;
            const uint32_t at = pairIndex[top - 1u - tid];
            pairGrad.ab[2 * uint64_t{at}] = acc[0];
            pairGrad.ab[2 * uint64_t{at} + 1] = acc[1];
            pairGrad.c[at] = acc8;
        }
        __syncthreads();
        top = lo;
    }
}

__global__ __launch_bounds__(kBlock) void blendBackwardKernel(FrameParams f, uint32_t slots,
                                                              const uint2* __restrict__ ranges,
                                                              const uint32_t* __restrict__ pairSlot,
                                                              const uint32_t* __restrict__ pairIndex,
                                                              const float4* __restrict__ proj,
                                                              const float4* __restrict__ accum,
                                                              const uint32_t* __restrict__ lastCount,
                                                              const float* __restrict__ finalT,
                                                              const float4* __restrict__ gradColour,
                                                              PairGrads pairGrad,
                                                              const uint32_t* __restrict__ order) {
    // Heaviest tiles first when an order is given: a long tile starts early
    // instead of running alone at the end.
    const uint32_t tile = order != nullptr ? order[blockIdx.x] : blockIdx.x;
    blendBackwardTile(f, tile, ranges[tile], slots, pairSlot, pairIndex, proj, accum, lastCount, finalT,
                      gradColour, pairGrad);
}

// Every tile of a batch of views: workgroup g takes batch tile g, in its view's
// frame and per-pixel arrays (ViewInfo.pixelBase), with that view's colour
// gradient.
__global__ __launch_bounds__(kBlock) void blendBackwardViewsKernel(const FrameParams* __restrict__ params,
                                                                   const ViewInfo* __restrict__ views,
                                                                   uint32_t viewCount, uint32_t slots,
                                                                   const uint2* __restrict__ ranges,
                                                                   const uint32_t* __restrict__ pairSlot,
                                                                   const uint32_t* __restrict__ pairIndex,
                                                                   const float4* __restrict__ proj,
                                                                   const float4* __restrict__ accum,
                                                                   const uint32_t* __restrict__ lastCount,
                                                                   const float* __restrict__ finalT,
                                                                   const float4* const* __restrict__ gradColours,
                                                                   PairGrads pairGrad,
                                                                   const uint32_t* __restrict__ order) {
    const uint32_t g = order != nullptr ? order[blockIdx.x] : blockIdx.x;
    const uint32_t v = viewOfTile(views, viewCount, g);
    const ViewInfo& info = views[v];
    const FrameParams f = params[info.params];
    blendBackwardTile(f, g - info.tileBase, ranges[g], slots, pairSlot, pairIndex, proj, accum + info.pixelBase,
                      lastCount + info.pixelBase, finalT + info.pixelBase, gradColours[v], pairGrad);
}

// ---- preprocess backward ------------------------------------------------------

// The rest of the colour and its gradients: given dL/d(rest) per channel,
// dL/d(coefficient) and dL/d(direction) (sh.slang:25-50 differentiated).
__device__ __forceinline__ void evaluateRestGrad(const float* k, uint32_t degree, float x, float y, float z,
                                                 const float dRest[3], float* dCoeff, float dDir[3]) {
    float basis[15];
    float grad[15][3];
    const float kC1 = 0.4886025119029199f;
    basis[0] = -kC1 * y; grad[0][0] = 0.0f; grad[0][1] = -kC1; grad[0][2] = 0.0f;
    basis[1] = kC1 * z;  grad[1][0] = 0.0f; grad[1][1] = 0.0f; grad[1][2] = kC1;
    basis[2] = -kC1 * x; grad[2][0] = -kC1; grad[2][1] = 0.0f; grad[2][2] = 0.0f;
    const float xx = x * x, yy = y * y, zz = z * z;
    const float c20 = 1.0925484305920792f, c21 = -1.0925484305920792f, c22 = 0.31539156525252005f,
                c23 = -1.0925484305920792f, c24 = 0.5462742152960396f;
    basis[3] = c20 * x * y;               grad[3][0] = c20 * y;  grad[3][1] = c20 * x;   grad[3][2] = 0.0f;
    basis[4] = c21 * y * z;               grad[4][0] = 0.0f;     grad[4][1] = c21 * z;   grad[4][2] = c21 * y;
    basis[5] = c22 * (2.0f * zz - xx - yy);
    grad[5][0] = -2.0f * c22 * x; grad[5][1] = -2.0f * c22 * y; grad[5][2] = 4.0f * c22 * z;
    basis[6] = c23 * x * z;               grad[6][0] = c23 * z;  grad[6][1] = 0.0f;      grad[6][2] = c23 * x;
    basis[7] = c24 * (xx - yy);           grad[7][0] = 2.0f * c24 * x; grad[7][1] = -2.0f * c24 * y; grad[7][2] = 0.0f;
    const float c30 = -0.5900435899266435f, c31 = 2.890611442640554f, c32 = -0.4570457994644658f,
                c33 = 0.3731763325901154f, c34 = -0.4570457994644658f, c35 = 1.445305721320277f,
                c36 = -0.5900435899266435f;
    basis[8] = c30 * y * (3.0f * xx - yy);
    grad[8][0] = c30 * 6.0f * x * y; grad[8][1] = c30 * (3.0f * xx - 3.0f * yy); grad[8][2] = 0.0f;
    basis[9] = c31 * x * y * z;
    grad[9][0] = c31 * y * z; grad[9][1] = c31 * x * z; grad[9][2] = c31 * x * y;
    basis[10] = c32 * y * (4.0f * zz - xx - yy);
    grad[10][0] = c32 * -2.0f * x * y; grad[10][1] = c32 * (4.0f * zz - xx - 3.0f * yy); grad[10][2] = c32 * 8.0f * y * z;
    basis[11] = c33 * z * (2.0f * zz - 3.0f * xx - 3.0f * yy);
    grad[11][0] = c33 * -6.0f * x * z; grad[11][1] = c33 * -6.0f * y * z;
    grad[11][2] = c33 * (6.0f * zz - 3.0f * xx - 3.0f * yy);
    basis[12] = c34 * x * (4.0f * zz - xx - yy);
    grad[12][0] = c34 * (4.0f * zz - 3.0f * xx - yy); grad[12][1] = c34 * -2.0f * x * y; grad[12][2] = c34 * 8.0f * x * z;
    basis[13] = c35 * z * (xx - yy);
    grad[13][0] = c35 * 2.0f * x * z; grad[13][1] = c35 * -2.0f * y * z; grad[13][2] = c35 * (xx - yy);
    basis[14] = c36 * x * (xx - 3.0f * yy);
    grad[14][0] = c36 * (3.0f * xx - 3.0f * yy); grad[14][1] = c36 * -6.0f * x * y; grad[14][2] = 0.0f;
    const uint32_t used = restCount(degree);
    dDir[0] = dDir[1] = dDir[2] = 0.0f;
#pragma unroll
    for (uint32_t bI = 0; bI < 15; ++bI) {
        const bool on = bI < used;
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) {
            dCoeff[bI * 3 + ch] = on ? basis[bI] * dRest[ch] : 0.0f;
            const float w = on ? k[bI * 3 + ch] * dRest[ch] : 0.0f;
            dDir[0] += w * grad[bI][0];
            dDir[1] += w * grad[bI][1];
            dDir[2] += w * grad[bI][2];
        }
    }
}

// A splat's parameter gradients, summed over the views that drew it. The
// harmonic entries past the cloud's own stride stay zero and are not written.
struct SplatGrad {
    float mean[3], scale[3], quat[4], logit;
    float sh[48];   // dc rgb, then 15 rest bases rgb
};

__device__ __forceinline__ void zero(SplatGrad& g) {
#pragma unroll
    for (int k = 0; k < 3; ++k) g.mean[k] = g.scale[k] = 0.0f;
#pragma unroll
    for (int k = 0; k < 4; ++k) g.quat[k] = 0.0f;
    g.logit = 0.0f;
#pragma unroll
    for (int k = 0; k < 48; ++k) g.sh[k] = 0.0f;
}

// The screen-space gradient g of splat i in frame f (d centre.xy, d conic a b
// c, d alpha, d rgb) walked back to its parameters and added into acc: conic
// -> 2D covariance (dilation, antialiasing factor), A = J W M, M = R diag(s),
// the clamped Jacobian, the centre, the rotation of the normalised
// quaternion, the 3D filter, exp and sigmoid, and the harmonics through the
// view direction.
__device__ __forceinline__ void splatBackward(const FrameParams& f, const DiffCloud& c, uint32_t i,
                                              const Decoded& dec, const float dc[3], const float coeff[45],
                                              const float g[9], SplatGrad& acc) {
    const float dCx = g[0], dCy = g[1], dCa = g[2], dCb = g[3], dCc = g[4], dAlpha = g[5];
    const float dCol[3] = {g[6], g[7], g[8]};

    // Recompute the forward (projectWave's footprint) and walk it back.
    const SplatShape& s = dec.shape;
    const float* m = f.m;
    const float tx = m[0] * s.px + m[1] * s.py + m[2] * s.pz + m[3];
    const float ty = m[4] * s.px + m[5] * s.py + m[6] * s.pz + m[7];
    const float tz = m[8] * s.px + m[9] * s.py + m[10] * s.pz + m[11];
    float r[9];
    rotationOf(s.q, r);
    const float sc[3] = {s.sx, s.sy, s.sz};
    float mm[9];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) mm[a * 3 + b] = r[a * 3 + b] * sc[b];
    float jx0, jx2, jy1, jy2;
    bool clampX = false, clampY = false;
    float limX = 0.0f, limY = 0.0f;
    const float invZ = 1.0f / tz;
    if (f.orthographic) {
        jx0 = f.focalX; jx2 = 0.0f;
        jy1 = f.focalY; jy2 = 0.0f;
    } else {
        const float limLoX = f.limLoX, limHiX = f.limHiX, limLoY = f.limLoY, limHiY = f.limHiY;
        const float ux = tx * invZ, uy = ty * invZ;
        clampX = ux < limLoX || ux > limHiX;
        clampY = uy < limLoY || uy > limHiY;
        limX = ux < limLoX ? limLoX : limHiX;
        limY = uy < limLoY ? limLoY : limHiY;
        const float ctx = clampS(ux, limLoX, limHiX) * tz;
        const float cty = clampS(uy, limLoY, limHiY) * tz;
        jx0 = f.focalX * invZ;
        jx2 = -f.focalX * ctx * invZ * invZ;
        jy1 = f.focalY * invZ;
        jy2 = -f.focalY * cty * invZ * invZ;
    }
    const float t3x[3] = {jx0 * m[0] + jx2 * m[8], jx0 * m[1] + jx2 * m[9], jx0 * m[2] + jx2 * m[10]};
    const float t3y[3] = {jy1 * m[4] + jy2 * m[8], jy1 * m[5] + jy2 * m[9], jy1 * m[6] + jy2 * m[10]};
    float ax[3], ay[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        ax[j] = t3x[0] * mm[0 * 3 + j] + t3x[1] * mm[1 * 3 + j] + t3x[2] * mm[2 * 3 + j];
        ay[j] = t3y[0] * mm[0 * 3 + j] + t3y[1] * mm[1 * 3 + j] + t3y[2] * mm[2 * 3 + j];
    }
    const float a0 = ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2];
    const float b = ax[0] * ay[0] + ax[1] * ay[1] + ax[2] * ay[2];
    const float c0 = ay[0] * ay[0] + ay[1] * ay[1] + ay[2] * ay[2];
    const float detB = a0 * c0 - b * b;
    const float a1 = a0 + f.dilation, c1 = c0 + f.dilation;
    const float det = a1 * c1 - b * b;
    const float invDet = 1.0f / det;

    // alpha = opacity * sqrt(max(detB, 0) / det) with antialiasing
    float dOpacity = dAlpha, dDet = 0.0f, dDetB = 0.0f;
    if (f.antialias) {
        const float ratio = fmaxS(detB, 0.0f) * invDet;
        const float root = __builtin_sqrtf(ratio);
        dOpacity = dAlpha * root;
        if (detB > 0.0f && root > 0.0f) {
            const float alpha = s.opacity * root;
            dDetB = dAlpha * alpha * 0.5f / detB;
            dDet = -dAlpha * alpha * 0.5f * invDet;
        }
    }
    // conic = (c1, -b, a1) / det
    dDet += -(dCa * c1 - dCb * b + dCc * a1) * invDet * invDet;
    float dA = dCc * invDet + dDet * c1 + dDetB * c0;
    float dC = dCa * invDet + dDet * a1 + dDetB * a0;
    const float dB = -dCb * invDet - 2.0f * b * (dDet + dDetB);
    // (a, b, c) = (ax.ax, ax.ay, ay.ay)
    float dAx[3], dAy[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        dAx[j] = 2.0f * dA * ax[j] + dB * ay[j];
        dAy[j] = dB * ax[j] + 2.0f * dC * ay[j];
    }
    // A = T3 M
    float dT3x[3], dT3y[3], dM[9];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        dT3x[a] = dAx[0] * mm[a * 3 + 0] + dAx[1] * mm[a * 3 + 1] + dAx[2] * mm[a * 3 + 2];
        dT3y[a] = dAy[0] * mm[a * 3 + 0] + dAy[1] * mm[a * 3 + 1] + dAy[2] * mm[a * 3 + 2];
#pragma unroll
        for (int j = 0; j < 3; ++j) dM[a * 3 + j] = t3x[a] * dAx[j] + t3y[a] * dAy[j];
    }
    // M = R diag(s)
    float dR[9], dS[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            dR[a * 3 + j] = dM[a * 3 + j] * sc[j];
            dS[j] += dM[a * 3 + j] * r[a * 3 + j];
        }
    // the 3D filter: s' = sqrt(s^2 + f^2), opacity * prod(s / s')
    float dSigmoid = dOpacity;
    float dRaw[3] = {dS[0], dS[1], dS[2]};
    if (f.useFilter3d) {
        const float fl = c.filter3d[i];
        const float factor = s.opacity / dec.sigmoid;
        dSigmoid = dOpacity * factor;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const float sj = dec.s[j], sp = sc[j];
            dRaw[j] = dS[j] * sj / sp + dOpacity * s.opacity * (fl * fl) / (sj * sp * sp);
        }
    }
#pragma unroll
    for (int j = 0; j < 3; ++j) acc.scale[j] = acc.scale[j] + dRaw[j] * dec.s[j];   // s = exp(log s)
    acc.logit = acc.logit + dSigmoid * dec.sigmoid * (1.0f - dec.sigmoid);

    // R(q), q normalised
    {
        const float x = s.q[0], y = s.q[1], z = s.q[2], w = s.q[3];
        const float dx = 2.0f * (dR[1] * y + dR[2] * z + dR[3] * y - 2.0f * dR[4] * x - dR[5] * w + dR[6] * z +
                                 dR[7] * w - 2.0f * dR[8] * x);
        const float dy = 2.0f * (-2.0f * dR[0] * y + dR[1] * x + dR[2] * w + dR[3] * x + dR[5] * z - dR[6] * w +
                                 dR[7] * z - 2.0f * dR[8] * y);
        const float dz = 2.0f * (-2.0f * dR[0] * z - dR[1] * w + dR[2] * x + dR[3] * w - 2.0f * dR[4] * z +
                                 dR[5] * y + dR[6] * x + dR[7] * y);
        const float dw = 2.0f * (-dR[1] * z + dR[2] * y + dR[3] * z - dR[5] * x - dR[6] * y + dR[7] * x);
        const float dq[4] = {dx, dy, dz, dw};
        if (dec.qLen > 0.0f) {
            const float proj1 = dq[0] * s.q[0] + dq[1] * s.q[1] + dq[2] * s.q[2] + dq[3] * s.q[3];
#pragma unroll
            for (int k = 0; k < 4; ++k) acc.quat[k] = acc.quat[k] + (dq[k] - s.q[k] * proj1) / dec.qLen;
        }
    }

    // J and the centre, from the view position t = W p + w
    float dT[3] = {0.0f, 0.0f, 0.0f};
    if (f.orthographic) {
        dT[0] = dCx * f.focalX;
        dT[1] = dCy * f.focalY;
    } else {
        const float dJx0 = dT3x[0] * m[0] + dT3x[1] * m[1] + dT3x[2] * m[2];
        const float dJx2 = dT3x[0] * m[8] + dT3x[1] * m[9] + dT3x[2] * m[10];
        const float dJy1 = dT3y[0] * m[4] + dT3y[1] * m[5] + dT3y[2] * m[6];
        const float dJy2 = dT3y[0] * m[8] + dT3y[1] * m[9] + dT3y[2] * m[10];
        const float iz2 = invZ * invZ;
        dT[2] += -(dJx0 * f.focalX + dJy1 * f.focalY) * iz2;
        if (clampX) {
            dT[2] += dJx2 * f.focalX * limX * iz2;
        } else {
            dT[0] += -dJx2 * f.focalX * iz2;
            dT[2] += dJx2 * 2.0f * f.focalX * tx * iz2 * invZ;
        }
        if (clampY) {
            dT[2] += dJy2 * f.focalY * limY * iz2;
        } else {
            dT[1] += -dJy2 * f.focalY * iz2;
            dT[2] += dJy2 * 2.0f * f.focalY * ty * iz2 * invZ;
        }
        dT[0] += dCx * f.focalX * invZ;
        dT[1] += dCy * f.focalY * invZ;
        dT[2] += -(dCx * f.focalX * tx + dCy * f.focalY * ty) * iz2;
    }
    float dP[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) dP[k] = m[0 + k] * dT[0] + m[4 + k] * dT[1] + m[8 + k] * dT[2];

    // colour = max(0.5 + C0 dc + rest(dir), 0), dir = (p - eye) / |p - eye|
    const uint32_t degree = shDegree(f);
    const float dxo = s.px - f.eye[0], dyo = s.py - f.eye[1], dzo = s.pz - f.eye[2];
    const float len = __builtin_sqrtf((dxo * dxo + dyo * dyo) + dzo * dzo);
    const float dir[3] = {dxo / len, dyo / len, dzo / len};
    float rest[3];
    evaluateRest(coeff, degree, dir[0], dir[1], dir[2], rest);
    float dPre[3];
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
        const float pre = (0.5f + kShC0 * dc[ch]) + rest[ch];
        dPre[ch] = pre > 0.0f ? dCol[ch] : 0.0f;
        acc.sh[ch] = acc.sh[ch] + kShC0 * dPre[ch];
    }
    float dCoeff[45], dDir[3];
    evaluateRestGrad(coeff, degree, dir[0], dir[1], dir[2], dPre, dCoeff, dDir);
#pragma unroll
    for (uint32_t k = 0; k < 45; ++k) acc.sh[3 + k] = acc.sh[3 + k] + dCoeff[k];
    const float along = dDir[0] * dir[0] + dDir[1] * dir[1] + dDir[2] * dir[2];
#pragma unroll
    for (int k = 0; k < 3; ++k) acc.mean[k] = acc.mean[k] + (dP[k] + (dDir[k] - dir[k] * along) / len);
}

// The screen-space gradients of a workgroup's kBlock consecutive slots, one a
// thread: each is the sum of its run of pair records, [first, first + n), in
// pair order. The slots' runs are consecutive in slot-major order, so together
// they are one range of records, [begin, end); the workgroup streams it through
// LDS kBlock records at a time with coalesced loads (the next chunk's loads in
// flight while this one is summed), and each thread adds the records of its own
// run from LDS. Every thread of the workgroup must call it (n = 0 for one with
// no slot); begin/end are the same for all.
struct PairChunk {
    float4 a[kBlock];
    float4 b[kBlock];
    float c[kBlock];
};

__device__ __forceinline__ void slotGradients(const PairGrads& pairGrad, uint32_t begin, uint32_t end,
                                              uint32_t first, uint32_t n, PairChunk& chunk, float g[9]) {
#pragma unroll
    for (int k = 0; k < 9; ++k) g[k] = 0.0f;
    const uint32_t tid = threadIdx.x;
    float4 na = make_float4(0.0f, 0.0f, 0.0f, 0.0f), nb = na;
    float nc = 0.0f;
    if (begin + tid < end) {
        const uint64_t at = begin + tid;
        na = pairGrad.ab[2 * at];
        nb = pairGrad.ab[2 * at + 1];
        nc = pairGrad.c[at];
    }
    for (uint32_t c0 = begin; c0 < end; c0 += kBlock) {
        chunk.a[tid] = na;
        chunk.b[tid] = nb;
        chunk.c[tid] = nc;
        __syncthreads();
        if (c0 + kBlock + tid < end) {
            const uint64_t at = c0 + kBlock + tid;
            na = pairGrad.ab[2 * at];
            nb = pairGrad.ab[2 * at + 1];
            nc = pairGrad.c[at];
        }
        const uint32_t lo = first > c0 ? first : c0;
        const uint32_t hi = first + n < c0 + kBlock ? first + n : c0 + kBlock;
        for (uint32_t q = lo; q < hi; ++q) {
            const float4 a = chunk.a[q - c0], b = chunk.b[q - c0];
            const float d = chunk.c[q - c0];
            g[0] += a.x; g[1] += a.y; g[2] += a.z; g[3] += a.w;
            g[4] += b.x; g[5] += b.y; g[6] += b.z; g[7] += b.w;
            g[8] += d;
        }
        __syncthreads();
    }
}

// The record range of a workgroup's slots: the first thread's run start and
// the last slot's run end (runs are consecutive, so that is every run).
__device__ __forceinline__ uint2 blockRunRange(uint32_t first, uint32_t n, bool lastSlot, uint32_t* sRange) {
    if (threadIdx.x == 0) sRange[0] = first;
    if (lastSlot) sRange[1] = first + n;
    __syncthreads();
    return make_uint2(sRange[0], sRange[1]);
}

// Every gradient but the harmonics' (those go through LDS, moveRows).
__device__ __forceinline__ void storeShapeGrad(const SplatGrad& acc, uint32_t i, const DiffGrads& out) {
    float* gMean = out.means + static_cast<uint64_t>(i) * 3;
    float* gScale = out.logScales + static_cast<uint64_t>(i) * 3;
    float* gQuat = out.quats + static_cast<uint64_t>(i) * 4;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        gMean[k] = acc.mean[k];
        gScale[k] = acc.scale[k];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) gQuat[k] = acc.quat[k];
    out.logits[i] = acc.logit;
}

__global__ __launch_bounds__(kBlock) void projectBackwardKernel(FrameParams f, DiffCloud c,
                                                                const uint32_t* __restrict__ touched,
                                                                const uint32_t* __restrict__ offsets,
                                                                PairGrads pairGrad, DiffGrads out) {
    // The pair chunks, then the workgroup's harmonic rows, then their gradients.
    __shared__ float sRows[kBlock * kRowPitchMax];
    __shared__ uint32_t sRange[2];
    static_assert(sizeof(PairChunk) <= sizeof(sRows), "the pair chunk fits the rows' LDS");
    PairChunk& chunk = *reinterpret_cast<PairChunk*>(sRows);
    const uint32_t blockFirst = blockIdx.x * kBlock;
    const uint32_t i = blockFirst + threadIdx.x;
    const bool live = i < f.count;
    const uint32_t rows = f.count - blockFirst < kBlock ? f.count - blockFirst : kBlock;
    const uint32_t slot = f.base + i;
    const uint32_t n = live ? touched[slot] : 0u;
    const uint32_t first = live ? offsets[slot] : 0u;
    const uint2 range = blockRunRange(first, n, threadIdx.x == rows - 1u, sRange);
    float g[9];
    slotGradients(pairGrad, range.x, range.y, first, n, chunk, g);
    __syncthreads();   // the chunk is done with
    const uint32_t stride = (1u + c.restPerColour) * 3u, pitch = rowPitch(stride);
    moveRows<true>(c.sh + static_cast<uint64_t>(blockFirst) * stride, stride, rows, sRows);
    __syncthreads();
    SplatGrad acc;
    zero(acc);
    if (n > 0) {
        // The coefficients read from the thread's LDS row (the degree's bases only).
        const Decoded dec = decode(c, i, f.useFilter3d != 0);
        const float* row = sRows + threadIdx.x * pitch;
        splatBackward(f, c, i, dec, row, row + 3, g, acc);
    }
    __syncthreads();   // every row is read
    if (live) {
        float* row = sRows + threadIdx.x * pitch;
#pragma unroll
        for (uint32_t k = 0; k < 48; ++k)
            if (k < stride) row[k] = acc.sh[k];
        storeShapeGrad(acc, i, out);
    }
    __syncthreads();
    moveRows<false>(out.sh + static_cast<uint64_t>(blockFirst) * stride, stride, rows, sRows);
}

// A batch of views of one cloud: each splat decodes its parameters once and
// adds its gradient in every view that drew it, in view order (no atomics).
// View v's frame is params[v]; the splat's slot there is params[v].base + i.
// The harmonic rows stay in LDS for every view (then hold their gradients), so
// the coefficients take no registers across the views.
__global__ __launch_bounds__(kBlock) void projectBackwardViewsKernel(const FrameParams* __restrict__ params,
                                                                     uint32_t views, DiffCloud c,
                                                                     const uint32_t* __restrict__ touched,
                                                                     const uint32_t* __restrict__ offsets,
                                                                     PairGrads pairGrad, DiffGrads out) {
    __shared__ float sRows[kBlock * kRowPitchMax];
    __shared__ PairChunk chunk;
    __shared__ uint32_t sRange[2];
    const uint32_t blockFirst = blockIdx.x * kBlock;
    const uint32_t i = blockFirst + threadIdx.x;
    const bool live = i < c.count;
    const uint32_t rows = c.count - blockFirst < kBlock ? c.count - blockFirst : kBlock;
    const FrameParams& f0 = params[0];
    const uint32_t stride = (1u + c.restPerColour) * 3u, pitch = rowPitch(stride);
    moveRows<true>(c.sh + static_cast<uint64_t>(blockFirst) * stride, stride, rows, sRows);
    Decoded dec{};
    if (live) dec = decode(c, i, f0.useFilter3d != 0);
    float* row = sRows + threadIdx.x * pitch;
    SplatGrad acc;
    zero(acc);
    __syncthreads();   // the rows are staged
    for (uint32_t v = 0; v < views; ++v) {
        const FrameParams& f = params[v];
        const uint32_t slot = f.base + i;
        const uint32_t n = live ? touched[slot] : 0u;
        const uint32_t first = live ? offsets[slot] : 0u;
        const uint2 range = blockRunRange(first, n, threadIdx.x == rows - 1u, sRange);
        float g[9];
        slotGradients(pairGrad, range.x, range.y, first, n, chunk, g);
        if (n > 0) splatBackward(f, c, i, dec, row, row + 3, g, acc);
        __syncthreads();   // sRange is rewritten for the next view
    }
    if (live) {
#pragma unroll
        for (uint32_t k = 0; k < 48; ++k)
            if (k < stride) row[k] = acc.sh[k];
        storeShapeGrad(acc, i, out);
    }
    __syncthreads();
    moveRows<false>(out.sh + static_cast<uint64_t>(blockFirst) * stride, stride, rows, sRows);
}

}   // namespace

void projectDiff(const FrameParams& f, const DiffCloud& cloud, float4* proj, uint2* rects, uint32_t* touched,
                 DepthBlocks depth, hipStream_t stream) {
    if (f.count == 0) return;
    projectDiffKernel<<<projectBlocks(f.count), kBlock, 0, stream>>>(f, cloud, proj, rects, touched, depth);
}

void projectDiffViews(const FrameParams* params, uint32_t views, const DiffCloud& cloud, float4* proj, uint2* rects,
                      uint32_t* touched, hipStream_t stream) {
    if (cloud.count == 0 || views == 0) return;
    projectDiffViewsKernel<<<(cloud.count + kBlock - 1) / kBlock, kBlock, 0, stream>>>(params, views, cloud, proj, rects,
                                                                                      touched);
}

void gatherSlots(const uint32_t* pairIndex, const uint32_t* pairSlot, uint32_t pairs, uint32_t* sortedSlots,
                 hipStream_t stream, const uint32_t* nDevice) {
    if (pairs == 0) return;
    gatherSlotsKernel<<<(pairs + kBlock - 1) / kBlock, kBlock, 0, stream>>>(pairIndex, pairSlot, pairs, nDevice,
                                                                            sortedSlots);
}

void blendBackward(const FrameParams& f, uint32_t slots, const uint2* ranges, const uint32_t* pairSlot,
                   const uint32_t* pairIndex, const float4* proj, const float4* accum, const uint32_t* lastCount,
                   const float* finalT, const float4* gradColour, PairGrads pairGrad, hipStream_t stream,
                   const uint32_t* order) {
    const uint32_t tiles = f.tilesX * f.tilesY;
    if (tiles == 0) return;
    blendBackwardKernel<<<tiles, kBlock, 0, stream>>>(f, slots, ranges, pairSlot, pairIndex, proj, accum,
                                                      lastCount, finalT, gradColour, pairGrad, order);
}

void blendBackwardViews(const FrameParams* params, const ViewInfo* views, uint32_t viewCount, uint32_t tiles,
                        uint32_t slots, const uint2* ranges, const uint32_t* pairSlot, const uint32_t* pairIndex,
                        const float4* proj, const float4* accum, const uint32_t* lastCount, const float* finalT,
                        const float4* const* gradColours, PairGrads pairGrad, hipStream_t stream,
                        const uint32_t* order) {
    if (tiles == 0 || viewCount == 0) return;
    blendBackwardViewsKernel<<<tiles, kBlock, 0, stream>>>(params, views, viewCount, slots, ranges, pairSlot,
                                                           pairIndex, proj, accum, lastCount, finalT, gradColours,
                                                           pairGrad, order);
}

void projectBackwardViews(const FrameParams* params, uint32_t views, const DiffCloud& cloud, const uint32_t* touched,
                          const uint32_t* offsets, PairGrads pairGrad, const DiffGrads& grads,
                          hipStream_t stream) {
    if (cloud.count == 0 || views == 0) return;
    projectBackwardViewsKernel<<<(cloud.count + kBlock - 1) / kBlock, kBlock, 0, stream>>>(params, views, cloud,
                                                                                          touched, offsets, pairGrad,
                                                                                          grads);
}

void projectBackward(const FrameParams& f, const DiffCloud& cloud, const uint32_t* touched, const uint32_t* offsets,
                     PairGrads pairGrad, const DiffGrads& grads, hipStream_t stream) {
    if (f.count == 0) return;
    projectBackwardKernel<<<(f.count + kBlock - 1) / kBlock, kBlock, 0, stream>>>(f, cloud, touched, offsets,
                                                                                  pairGrad, grads);
}

}   // namespace lrt::kern
