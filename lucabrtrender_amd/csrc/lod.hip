// Levels of detail on the GPU: see lod.h. The method is the reference's
// (shaders/lrt/lod/lod_common.slang:1-33); the launches are this build's own:
//   build  Morton keys -> the library's onesweep sort (30 bits) -> the splats
//          gathered into that order -> one pass giving each splat the coarsest
//          level at which it starts a cell (so every level's group count comes
//          from one histogram and one read) -> per stored level, finest first,
//          one scan of "starts a cell at this level" giving groups, starts and
//          cell codes -> moments (the finest level from its splats, each
//          coarser one as sums of its children) -> one Gaussian a group.
//   cut    one flag kernel over every level's groups and every splat, one
//          scan over all of them, one read of the two counts, one gather:
//          the reference's per-level cut, scan and gather dispatches
//          (Lod.cpp:440-525) in four launches, in the same output order
//          (levels coarsest first, then the splats in Morton order).
// The moments sum each group's splats in order in FP32 and the finalize is the
// reference's cyclic Jacobi, op for op: the CPU restatement reproduces every
// merged Gaussian bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "kernels.h"
#include "lod.h"
#include "packing.h"
#include "projection.h"
#include "splat_math.h"

namespace lrt::lod {
namespace {

using namespace lrt::dev;

constexpr uint32_t kBlock = 256;
constexpr uint32_t kBins = kLevels + 2;   // a splat starts a cell from level 0 (the first) .. 10, or never (11)

// ---- Morton codes (shaders/lrt/common/morton.slang:7-34) ---------------------

__device__ __forceinline__ uint32_t spreadBits(uint32_t v) {
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__device__ __forceinline__ uint32_t compactBits(uint32_t v) {
    v &= 0x09249249u;
    v = (v | (v >> 2)) & 0x030C30C3u;
    v = (v | (v >> 4)) & 0x0300F00Fu;
    v = (v | (v >> 8)) & 0x030000FFu;
    v = (v | (v >> 16)) & 0x000003FFu;
    return v;
}

__device__ __forceinline__ uint32_t quantise(float unit) {
    const float s = unit > 0.0f ? (unit < 1.0f ? unit : 1.0f) : 0.0f;   // saturate, NaN to 0
    const uint32_t q = static_cast<uint32_t>(s * 1024.0f);
    return q < 1023u ? q : 1023u;
}

struct Bounds {
    float lo[3];
    float extent;
};

__global__ __launch_bounds__(kBlock) void mortonKernel(const float4* __restrict__ positions, uint32_t n, Bounds b,
                                                       uint64_t* __restrict__ keys, uint32_t* __restrict__ values) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const float4 p = positions[i];
    const float e = fmaxS(b.extent, 1e-20f);
    const uint32_t code = (spreadBits(quantise((p.x - b.lo[0]) / e)) << 2) |
                          (spreadBits(quantise((p.y - b.lo[1]) / e)) << 1) | spreadBits(quantise((p.z - b.lo[2]) / e));
    keys[i] = code;
    values[i] = i;
}

// Slot k takes splat order[k]: the cloud in Morton order (lod_reorder.slang).
__global__ __launch_bounds__(kBlock) void reorderKernel(const uint32_t* __restrict__ order, uint32_t n,
                                                        const float4* __restrict__ srcPositions,
                                                        const uint4* __restrict__ srcShape,
                                                        const uint32_t* __restrict__ srcSh, uint32_t shWords,
                                                        float4* __restrict__ positions, uint4* __restrict__ shape,
                                                        uint32_t* __restrict__ sh) {
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= n) return;
    const uint32_t i = order[k];
    positions[k] = srcPositions[i];
    shape[k] = srcShape[i];
    if (sh != nullptr) {
        for (uint32_t w = 0; w < shWords; ++w) {
            sh[static_cast<uint64_t>(k) * shWords + w] = srcSh[static_cast<uint64_t>(i) * shWords + w];
        }
    }
}

// The coarsest level at which splat i starts a cell: its code's top 3r bits
// differ from splat i - 1's exactly for r past a third of the codes' common
// leading bits (lod_boundaries.slang's test at every level at once). The first
// splat starts one at every level (0); a code equal to its predecessor's never
// does (11). The histogram of those levels gives every level's group count.
__global__ __launch_bounds__(kBlock) void firstLevelKernel(const uint64_t* __restrict__ keys, uint32_t n,
                                                           uint8_t* __restrict__ first, uint32_t* __restrict__ hist) {
    __shared__ uint32_t sHist[kBins];
    if (threadIdx.x < kBins) sHist[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) {
        uint32_t level = 0;
        if (i > 0) {
            const uint32_t x = static_cast<uint32_t>(keys[i] ^ keys[i - 1]);
            // common leading bits of the 30-bit codes: clz(x) - 2
            level = x == 0 ? kLevels + 1 : (static_cast<uint32_t>(__clz(x)) - 2u) / 3u + 1u;
        }
        first[i] = static_cast<uint8_t>(level);
        atomicAdd(&sHist[level], 1u);
    }
    __syncthreads();
    if (threadIdx.x < kBins && sHist[threadIdx.x] != 0) atomicAdd(&hist[threadIdx.x], sHist[threadIdx.x]);
}

__global__ __launch_bounds__(kBlock) void startsCellKernel(const uint8_t* __restrict__ first, uint32_t n,
                                                           uint32_t level, uint32_t* __restrict__ flag) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) flag[i] = first[i] <= level ? 1u : 0u;
}

// Every splat's group at `level`, every group's first splat and cell code
// (lod_groups.slang): the group that starts at a splat is written by it.
__global__ __launch_bounds__(kBlock) void groupsKernel(const uint64_t* __restrict__ keys,
                                                       const uint32_t* __restrict__ flag,
                                                       const uint32_t* __restrict__ before, uint32_t n,
                                                       uint32_t level, uint32_t* __restrict__ group,
                                                       uint32_t* __restrict__ starts, uint32_t* __restrict__ cells) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t g = before[i] + flag[i] - 1u;
    group[i] = g;
    if (flag[i] != 0) {
        starts[g] = i;
        cells[g] = static_cast<uint32_t>(keys[i]) >> (3u * (kLevels - level));
    }
}

// ---- moments (lod_leaf_moments.slang, lod_merge_moments.slang) ---------------

__global__ __launch_bounds__(64) void leafMomentsKernel(const float4* __restrict__ positions,
                                                        const uint4* __restrict__ shape,
                                                        const uint32_t* __restrict__ sh, uint32_t n,
                                                        const uint32_t* __restrict__ starts, uint32_t groups,
                                                        uint32_t keep, uint32_t shWords, uint32_t stride,
                                                        float* __restrict__ moments) {
    const uint32_t g = blockIdx.x * 64 + threadIdx.x;
    if (g >= groups) return;
    const uint32_t first = starts[g];
    const uint32_t end = g + 1 < groups ? starts[g + 1] : n;
    float m[kMomentsHead] = {};
    float rest[45] = {};
    for (uint32_t i = first; i < end; ++i) {
        const float4 p = positions[i];
        const uint4 row = shape[i];
        const float s[3] = {expF(lowHalf(row.y)), expF(highHalf(row.y)), expF(lowHalf(row.z))};
        float q[4];
        decodeQuaternion(row.x, q);
        float r[9];
        rotationOf(q, r);
        float mm[9];
        for (int a = 0; a < 3; ++a)
            for (int c = 0; c < 3; ++c) mm[a * 3 + c] = r[a * 3 + c] * s[c];
        float cov[9];
        for (int a = 0; a < 3; ++a)
            for (int c = 0; c < 3; ++c)
                cov[a * 3 + c] = (mm[a * 3] * mm[c * 3] + mm[a * 3 + 1] * mm[c * 3 + 1]) + mm[a * 3 + 2] * mm[c * 3 + 2];
        const float smallest = fminS(fminS(s[0], s[1]), s[2]);
        const float area = s[0] * s[1] * s[2] / fmaxS(smallest, 1e-20f);
        const float w = p.w * area;
        m[0] += w;
        m[1] += w * p.x;
        m[2] += w * p.y;
        m[3] += w * p.z;
        m[4] += w * (cov[0] + p.x * p.x);
        m[5] += w * (cov[1] + p.x * p.y);
        m[6] += w * (cov[2] + p.x * p.z);
        m[7] += w * (cov[4] + p.y * p.y);
        m[8] += w * (cov[5] + p.y * p.z);
        m[9] += w * (cov[8] + p.z * p.z);
        m[10] += w * highHalf(row.z);
        m[11] += w * lowHalf(row.w);
        m[12] += w * highHalf(row.w);
        for (uint32_t h = 0; h < keep * 3; ++h) {
            const uint32_t word = sh[static_cast<uint64_t>(i) * shWords + h / 2];
            rest[h] += w * ((h & 1u) == 0u ? lowHalf(word) : highHalf(word));
        }
    }
    float* out = moments + static_cast<uint64_t>(g) * stride;
    for (uint32_t k = 0; k < kMomentsHead; ++k) out[k] = m[k];
    for (uint32_t h = 0; h < keep * 3; ++h) out[kMomentsHead + h] = rest[h];
}

// A coarser group's moments: its children's, summed in order. Its children are
// the finer groups from that of its first splat to that of the next group's.
__global__ __launch_bounds__(64) void mergeMomentsKernel(const uint32_t* __restrict__ starts, uint32_t groups,
                                                         const uint32_t* __restrict__ fineGroup, uint32_t fineGroups,
                                                         const float* __restrict__ fineMoments, uint32_t stride,
                                                         float* __restrict__ moments) {
    const uint32_t g = blockIdx.x * 64 + threadIdx.x;
    if (g >= groups) return;
    const uint32_t first = fineGroup[starts[g]];
    const uint32_t end = g + 1 < groups ? fineGroup[starts[g + 1]] : fineGroups;
    for (uint32_t k = 0; k < stride; ++k) {
        float sum = 0.0f;
        for (uint32_t h = first; h < end; ++h) sum += fineMoments[static_cast<uint64_t>(h) * stride + k];
        moments[static_cast<uint64_t>(g) * stride + k] = sum;
    }
}

// ---- one Gaussian a group (lod_finalize.slang) ------------------------------

struct M3 {
    float a[3][3];
};

__device__ __forceinline__ M3 mul3(const M3& x, const M3& y) {
    M3 r;
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) r.a[i][k] = (x.a[i][0] * y.a[0][k] + x.a[i][1] * y.a[1][k]) + x.a[i][2] * y.a[2][k];
    return r;
}

__device__ __forceinline__ M3 transposed(const M3& x) {
    M3 r;
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) r.a[i][k] = x.a[k][i];
    return r;
}

__device__ __forceinline__ float signOf(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }

// Cyclic Jacobi on a symmetric 3x3, six sweeps: `a` ends diagonal and v's
// columns are the eigenvectors.
__device__ __forceinline__ void jacobi(M3& a, M3& v) {
    v = M3{{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}};
    for (int sweep = 0; sweep < 6; ++sweep) {
        for (int pq = 0; pq < 3; ++pq) {
            const int p = pq == 2 ? 1 : 0;
            const int q = pq == 0 ? 1 : 2;
            const float apq = a.a[p][q];
            if (__builtin_fabsf(apq) < 1e-30f) continue;
            const float theta = (a.a[q][q] - a.a[p][p]) / (2.0f * apq);
            const float t = signOf(theta + 1e-30f) / (__builtin_fabsf(theta) + __builtin_sqrtf(theta * theta + 1.0f));
            const float c = 1.0f / __builtin_sqrtf(t * t + 1.0f);
            const float s = t * c;
            M3 j{{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}};
            j.a[p][p] = c;
            j.a[q][q] = c;
            j.a[p][q] = s;
            j.a[q][p] = -s;
            a = mul3(transposed(j), mul3(a, j));
            v = mul3(v, j);
        }
    }
}

// The unit quaternion (x, y, z, w) of a rotation matrix (packing.slang:76-93).
__device__ __forceinline__ void quaternionOfAxes(const M3& m, float q[4]) {
    const float trace = m.a[0][0] + m.a[1][1] + m.a[2][2];
    if (trace > 0.0f) {
        const float s = 0.5f / __builtin_sqrtf(trace + 1.0f);
        q[0] = (m.a[2][1] - m.a[1][2]) * s;
        q[1] = (m.a[0][2] - m.a[2][0]) * s;
        q[2] = (m.a[1][0] - m.a[0][1]) * s;
        q[3] = 0.25f / s;
    } else if (m.a[0][0] > m.a[1][1] && m.a[0][0] > m.a[2][2]) {
        const float s = 2.0f * __builtin_sqrtf(1.0f + m.a[0][0] - m.a[1][1] - m.a[2][2]);
        q[0] = 0.25f * s;
        q[1] = (m.a[1][0] + m.a[0][1]) / s;
        q[2] = (m.a[0][2] + m.a[2][0]) / s;
        q[3] = (m.a[2][1] - m.a[1][2]) / s;
    } else if (m.a[1][1] > m.a[2][2]) {
        const float s = 2.0f * __builtin_sqrtf(1.0f - m.a[0][0] + m.a[1][1] - m.a[2][2]);
        q[0] = (m.a[1][0] + m.a[0][1]) / s;
        q[1] = 0.25f * s;
        q[2] = (m.a[2][1] + m.a[1][2]) / s;
        q[3] = (m.a[0][2] - m.a[2][0]) / s;
    } else {
        const float s = 2.0f * __builtin_sqrtf(1.0f - m.a[0][0] - m.a[1][1] + m.a[2][2]);
        q[0] = (m.a[0][2] + m.a[2][0]) / s;
        q[1] = (m.a[2][1] + m.a[1][2]) / s;
        q[2] = 0.25f * s;
        q[3] = (m.a[1][0] - m.a[0][1]) / s;
    }
    const float len = __builtin_sqrtf(((q[0] * q[0] + q[1] * q[1]) + q[2] * q[2]) + q[3] * q[3]);
    for (int k = 0; k < 4; ++k) q[k] = q[k] / len;
}

__global__ __launch_bounds__(64) void finalizeKernel(const float* __restrict__ moments, uint32_t groups,
                                                     uint32_t keep, uint32_t shWords, uint32_t stride,
                                                     float4* __restrict__ positions, uint4* __restrict__ shape,
                                                     uint32_t* __restrict__ sh) {
    const uint32_t g = blockIdx.x * 64 + threadIdx.x;
    if (g >= groups) return;
    const float* m = moments + static_cast<uint64_t>(g) * stride;
    const float weight = fmaxS(m[0], 1e-30f);
    const float mu[3] = {m[1] / weight, m[2] / weight, m[3] / weight};
    M3 cov{{{m[4] / weight, m[5] / weight, m[6] / weight},
            {m[5] / weight, m[7] / weight, m[8] / weight},
            {m[6] / weight, m[8] / weight, m[9] / weight}}};
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) cov.a[i][k] = cov.a[i][k] - mu[i] * mu[k];
    M3 axes;
    jacobi(cov, axes);
    const float det = axes.a[0][0] * (axes.a[1][1] * axes.a[2][2] - axes.a[1][2] * axes.a[2][1]) -
                      axes.a[0][1] * (axes.a[1][0] * axes.a[2][2] - axes.a[1][2] * axes.a[2][0]) +
                      axes.a[0][2] * (axes.a[1][0] * axes.a[2][1] - axes.a[1][1] * axes.a[2][0]);
    if (det < 0.0f) {
        axes.a[0][2] = -axes.a[0][2];
        axes.a[1][2] = -axes.a[1][2];
        axes.a[2][2] = -axes.a[2][2];
    }
    const float s[3] = {__builtin_sqrtf(fmaxS(cov.a[0][0], 1e-14f)), __builtin_sqrtf(fmaxS(cov.a[1][1], 1e-14f)),
                        __builtin_sqrtf(fmaxS(cov.a[2][2], 1e-14f))};
    const float smallest = fminS(fminS(s[0], s[1]), s[2]);
    const float area = s[0] * s[1] * s[2] / fmaxS(smallest, 1e-20f);
    const float opacity = fminS(m[0] / fmaxS(area, 1e-30f), 0.99f);
    positions[g] = make_float4(mu[0], mu[1], mu[2], opacity);
    float q[4];
    quaternionOfAxes(axes, q);
    shape[g] = make_uint4(pack::encodeQuaternion(q), pack::packHalves(logF(s[0]), logF(s[1])),
                          pack::packHalves(logF(s[2]), m[10] / weight),
                          pack::packHalves(m[11] / weight, m[12] / weight));
    if (keep == 0) return;
    uint32_t* row = sh + static_cast<uint64_t>(g) * shWords;
    for (uint32_t h = 0; h < keep * 3; h += 2) {
        const float lo = m[kMomentsHead + h] / weight;
        const float hi = h + 1 < keep * 3 ? m[kMomentsHead + h + 1] / weight : 0.0f;
        row[h / 2] = h + 1 < keep * 3 ? pack::packHalves(lo, hi)
                                       : static_cast<uint32_t>(pack::floatToHalf(pack::halfSafe(lo)));
    }
}

// ---- the cut (lod_cut.slang, lod_gather.slang) -------------------------------

struct CutParams {
    uint32_t merged, splats, levels, finestFirst, finestLevel;
    uint32_t levelOf[kLevels];
    uint32_t first[kLevels + 1];
    float lo[3], extent, eye[3], pixelsPerUnit, threshold;
    uint32_t orthographic;
};

// The pixels a level-`level` cell spans at its nearest point to the eye.
__device__ __forceinline__ float projectedEdge(const CutParams& p, uint32_t cell, uint32_t level) {
    const float edge = p.extent / static_cast<float>(1u << level);
    if (p.orthographic) return edge * p.pixelsPerUnit;
    const uint32_t ix = compactBits(cell >> 2), iy = compactBits(cell >> 1), iz = compactBits(cell);
    const float lo[3] = {p.lo[0] + static_cast<float>(ix) * edge, p.lo[1] + static_cast<float>(iy) * edge,
                         p.lo[2] + static_cast<float>(iz) * edge};
    float d[3];
    for (int k = 0; k < 3; ++k) d[k] = fminS(fmaxS(p.eye[k], lo[k]), lo[k] + edge) - p.eye[k];
    const float distance = fmaxS(__builtin_sqrtf((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]), 1e-6f);
    return edge * p.pixelsPerUnit / distance;
}

__device__ __forceinline__ bool smallCell(const CutParams& p, uint32_t cell, uint32_t level) {
    return p.threshold > 0.0f && projectedEdge(p, cell, level) <= p.threshold;
}

// Element e < merged is a merged Gaussian (drawn when its cell is small and its
// parent's is not); element merged + i is splat i (drawn when its finest-level
// cell is not small).
__global__ __launch_bounds__(kBlock) void cutKernel(CutParams p, const uint32_t* __restrict__ cells,
                                                    const uint32_t* __restrict__ groups,
                                                    uint32_t* __restrict__ selected) {
    const uint32_t e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= p.merged + p.splats) return;
    bool drawn;
    if (e < p.merged) {
        uint32_t k = 0;
        while (k + 1 < p.levels && e >= p.first[k + 1]) ++k;
        const uint32_t level = p.levelOf[k];
        const uint32_t cell = cells[e];
        const bool here = smallCell(p, cell, level);
        const bool above = k > 0 && smallCell(p, cell >> 3, level - 1);
        drawn = here && !above;
    } else {
        const uint32_t g = groups[e - p.merged];
        drawn = !smallCell(p, cells[p.finestFirst + g], p.finestLevel);
    }
    selected[e] = drawn ? 1u : 0u;
}

// The drawn elements of a block of kBlock, compacted: they keep their order,
// so their destinations are consecutive from the block's first (the exclusive
// sum at its first element). The block lists them in LDS, copies their
// positions and shapes one thread an element, then their harmonic rows one
// word a thread across the block, so both sides of the row copy are runs of
// consecutive words.
__global__ __launch_bounds__(kBlock) void gatherKernel(uint32_t merged, uint32_t total,
                                                       const uint32_t* __restrict__ selected,
                                                       const uint32_t* __restrict__ dest,
                                                       const float4* __restrict__ mergedPositions,
                                                       const uint4* __restrict__ mergedShape,
                                                       const uint32_t* __restrict__ mergedSh,
                                                       const float4* __restrict__ splatPositions,
                                                       const uint4* __restrict__ splatShape,
                                                       const uint32_t* __restrict__ splatSh, uint32_t shWords,
                                                       float4* __restrict__ positions, uint4* __restrict__ shape,
                                                       uint32_t* __restrict__ sh) {
    __shared__ uint32_t sElement[kBlock];
    const uint32_t first = blockIdx.x * kBlock;
    const uint32_t last = first + kBlock - 1u < total ? first + kBlock - 1u : total - 1u;
    const uint32_t base = dest[first];
    const uint32_t count = dest[last] + selected[last] - base;
    const uint32_t e = first + threadIdx.x;
    if (e < total && selected[e] != 0) sElement[dest[e] - base] = e;
    __syncthreads();
    if (threadIdx.x < count) {
        const uint32_t k = sElement[threadIdx.x];
        const bool isMerged = k < merged;
        positions[base + threadIdx.x] = isMerged ? mergedPositions[k] : splatPositions[k - merged];
        shape[base + threadIdx.x] = isMerged ? mergedShape[k] : splatShape[k - merged];
    }
    if (sh == nullptr) return;
    const uint32_t words = count * shWords;
    uint32_t row = threadIdx.x / shWords, col = threadIdx.x - row * shWords;
    const uint32_t rowStep = kBlock / shWords, colStep = kBlock - rowStep * shWords;
    uint32_t* out = sh + static_cast<uint64_t>(base) * shWords;
    constexpr uint32_t kBatch = 8;   // loads in flight before the first store
    for (uint32_t w0 = threadIdx.x; w0 < words; w0 += kBlock * kBatch) {
        uint32_t v[kBatch];
#pragma unroll
        for (uint32_t b = 0; b < kBatch; ++b) {
            if (w0 + b * kBlock < words) {
                const uint32_t k = sElement[row];
                const uint32_t* src = k < merged ? mergedSh + static_cast<uint64_t>(k) * shWords
                                                 : splatSh + static_cast<uint64_t>(k - merged) * shWords;
                v[b] = src[col];
            }
            row += rowStep;
            col += colStep;
            if (col >= shWords) {
                col -= shWords;
                row += 1;
            }
        }
#pragma unroll
        for (uint32_t b = 0; b < kBatch; ++b) {
            if (w0 + b * kBlock < words) out[w0 + b * kBlock] = v[b];
        }
    }
}

uint32_t blocksFor(uint64_t n, uint32_t block) { return static_cast<uint32_t>((n + block - 1) / block); }

Status check(hipError_t e, const char* what) {
    if (e == hipSuccess) return {};
    return Status::error(e == hipErrorOutOfMemory ? ErrorCode::OutOfMemory : ErrorCode::DeviceFailure,
                         std::string(what) + ": " + hipGetErrorString(e));
}

}   // namespace

SplatCloud LodCloud::splats() const {
    SplatCloud c;
    c.positions = positions.as<float4>();
    c.shape = shape.as<uint4>();
    c.sh = restPerColour > 0 ? sh.as<uint32_t>() : nullptr;
    c.count = count;
    c.restPerColour = restPerColour;
    c.shWords = shWords;
    return c;
}

SplatCloud LodCloud::gaussians(size_t level) const {
    const LodLevel& l = levels.at(level);
    SplatCloud c;
    c.positions = mergedPositions.as<float4>() + l.first;
    c.shape = mergedShape.as<uint4>() + l.first;
    c.sh = restPerColour > 0 ? mergedSh.as<uint32_t>() + uint64_t{l.first} * shWords : nullptr;
    c.count = l.count;
    c.restPerColour = restPerColour;
    c.shWords = shWords;
    return c;
}

Status build(int device, hipStream_t stream, const SplatCloud& cloud, const LodBuildSettings& settings,
             LodCloud& out) {
    const uint32_t n = cloud.count;
    if (n == 0) return Status::error(ErrorCode::InvalidArgument, "an empty cloud has no levels of detail");
    if (cloud.positions == nullptr || cloud.shape == nullptr) {
        return Status::error(ErrorCode::InvalidArgument, "a cloud without positions or shape");
    }
    if (cloud.restPerColour != 0 && cloud.restPerColour != 3 && cloud.restPerColour != 8 &&
        cloud.restPerColour != 15) {
        return Status::error(ErrorCode::InvalidArgument, "restPerColour must be 0, 3, 8 or 15");
    }
    if (cloud.restPerColour > 0 && (cloud.sh == nullptr || cloud.shWords < (cloud.restPerColour * 3 + 1) / 2)) {
        return Status::error(ErrorCode::InvalidArgument, "a cloud with harmonics but no (or too short) sh rows");
    }
    if (n > (1u << 30) - 1u) return Status::error(ErrorCode::Unsupported, "more than 2^30 - 1 splats");
    Status s = check(hipSetDevice(device), "hipSetDevice");
    if (!s.ok()) return s;
    const uint32_t keep = cloud.restPerColour;
    const uint32_t shWords = keep > 0 ? cloud.shWords : 1u;
    out.levels.clear();
    out.extent = 0.0F;
    out.count = n;
    out.restPerColour = keep;
    out.shWords = shWords;

    // Scratch: bounds, keys and values (ping-pong), the sort's and the scan's
    // workspaces, per-splat level and flags, and the small counters.
    DeviceBuffer partial, keys[2], values[2], work, first, flag, before, scanStatus, counters;
    const uint64_t sortWords = kern::sortWorkspaceWords(n, 30);
    if (!(s = partial.reserve(uint64_t{kern::boundsPartials(n)} * 6 * 4 + 64, "lod.bounds")).ok()) return s;
    for (int k = 0; k < 2; ++k) {
        if (!(s = keys[k].reserve(uint64_t{n} * 8, "lod.keys")).ok()) return s;
        if (!(s = values[k].reserve(uint64_t{n} * 4, "lod.order")).ok()) return s;
    }
    if (!(s = work.reserve(sortWords * 4, "lod.sortWorkspace")).ok()) return s;
    if (!(s = first.reserve(n, "lod.first")).ok()) return s;
    if (!(s = flag.reserve(uint64_t{n} * 4, "lod.flag")).ok()) return s;
    if (!(s = before.reserve(uint64_t{n} * 4, "lod.before")).ok()) return s;
    if (!(s = scanStatus.reserve(kern::scanStatusWords(n) * 8, "lod.scanStatus")).ok()) return s;
    if (!(s = counters.reserve(64 * 4, "lod.counters")).ok()) return s;
    uint32_t* cnt = counters.as<uint32_t>();   // [0, 12) histogram, 16 sort error, 17 scan total, 18 nonzero, 19 scan error

    // The bounds, as the loader's (the extent of the positions).
    float* box = partial.as<float>() + uint64_t{kern::boundsPartials(n)} * 6;
    kern::splatBounds(cloud.positions, n, partial.as<float>(), box, stream);
    float hostBox[6];
    if (!(s = check(hipMemcpyAsync(hostBox, box, sizeof hostBox, hipMemcpyDeviceToHost, stream), "bounds")).ok())
        return s;
    if (!(s = check(hipStreamSynchronize(stream), "bounds")).ok()) return s;
    for (int k = 0; k < 3; ++k) {
        out.boundsLo[k] = hostBox[k];
        out.extent = std::max(out.extent, hostBox[3 + k] - hostBox[k]);
    }
    out.extent = std::max(out.extent, 1e-6F);
    out.extent *= 1.0001F;   // the far corner quantises inside the last cell

    // Morton order.
    Bounds b{{out.boundsLo[0], out.boundsLo[1], out.boundsLo[2]}, out.extent};
    mortonKernel<<<blocksFor(n, kBlock), kBlock, 0, stream>>>(cloud.positions, n, b, keys[0].as<uint64_t>(),
                                                              values[0].as<uint32_t>());
    kern::fill32(cnt, 0u, 64, stream);
    const kern::SortWorkspace ws = kern::sortWorkspace(work.as<uint32_t>(), n, 30, cnt + 16);
    const int sorted = kern::sortPairs(keys[0].as<uint64_t>(), values[0].as<uint32_t>(), keys[1].as<uint64_t>(),
                                       values[1].as<uint32_t>(), n, kern::KeyMode::Plain, 0, 30, ws, stream)
                           ? 1
                           : 0;
    const uint64_t* code = keys[sorted].as<uint64_t>();
    if (!(s = out.positions.reserve(uint64_t{n} * 16, "lod.positions")).ok()) return s;
    if (!(s = out.shape.reserve(uint64_t{n} * 16, "lod.shape")).ok()) return s;
    if (!(s = out.sh.reserve(uint64_t{n} * shWords * 4, "lod.sh")).ok()) return s;
    reorderKernel<<<blocksFor(n, kBlock), kBlock, 0, stream>>>(
        values[sorted].as<uint32_t>(), n, cloud.positions, cloud.shape, cloud.sh, shWords, out.positions.as<float4>(),
        out.shape.as<uint4>(), keep > 0 ? out.sh.as<uint32_t>() : nullptr);

    // Every level's group count from one histogram.
    firstLevelKernel<<<blocksFor(n, kBlock), kBlock, 0, stream>>>(code, n, first.as<uint8_t>(), cnt);
    uint32_t hostCnt[20];
    if (!(s = check(hipMemcpyAsync(hostCnt, cnt, sizeof hostCnt, hipMemcpyDeviceToHost, stream), "levels")).ok())
        return s;
    if (!(s = check(hipStreamSynchronize(stream), "levels")).ok()) return s;
    if (hostCnt[16] != 0) return Status::error(ErrorCode::InternalError, "the sort's look-back gave up");
    std::array<uint32_t, kLevels + 1> groupsAt{};
    uint32_t running = 0;
    for (uint32_t r = 0; r <= kLevels; ++r) {
        running += hostCnt[r];
        groupsAt[r] = running;
    }
    // The finest level worth storing (Lod.cpp:187-198): coarse to fine, until
    // the cells hold fewer than 1 / fraction splats on average.
    const uint32_t coarsest = std::clamp<uint32_t>(settings.coarsestLevel, 1, kLevels);
    uint32_t finest = coarsest;
    for (uint32_t r = coarsest; r <= kLevels; ++r) {
        if (static_cast<float>(groupsAt[r]) > settings.maxGroupFraction * static_cast<float>(n) && r > coarsest) break;
        finest = r;
    }
    uint32_t merged = 0;
    for (uint32_t r = coarsest; r <= finest; ++r) {
        out.levels.push_back(LodLevel{r, merged, groupsAt[r]});
        merged += groupsAt[r];
    }
    out.merged = merged;
    if (!(s = out.mergedPositions.reserve(uint64_t{merged} * 16, "lod.merged")).ok()) return s;
    if (!(s = out.mergedShape.reserve(uint64_t{merged} * 16, "lod.merged")).ok()) return s;
    if (!(s = out.mergedSh.reserve(uint64_t{merged} * shWords * 4, "lod.merged")).ok()) return s;
    if (!(s = out.cells.reserve(uint64_t{merged} * 4, "lod.cells")).ok()) return s;

    // Finest level first: groups, moments, Gaussians; each coarser level's
    // moments as its children's sums.
    const uint32_t stride = kMomentsHead + keep * 3;
    DeviceBuffer group[2], starts[2], moments[2];
    int cur = 0;
    for (uint32_t r = finest + 1; r-- > coarsest;) {
        const LodLevel& level = out.levels[r - coarsest];
        const uint32_t groups = level.count;
        if (!(s = group[cur].reserve(uint64_t{n} * 4, "lod.group")).ok()) return s;
        if (!(s = starts[cur].reserve(uint64_t{groups} * 4, "lod.starts")).ok()) return s;
        if (!(s = moments[cur].reserve(uint64_t{groups} * stride * 4, "lod.moments")).ok()) return s;
        startsCellKernel<<<blocksFor(n, kBlock), kBlock, 0, stream>>>(first.as<uint8_t>(), n, r, flag.as<uint32_t>());
        kern::scanExclusive(flag.as<uint32_t>(), before.as<uint32_t>(), n, scanStatus.as<uint64_t>(), cnt + 17,
                            cnt + 18, cnt + 19, stream);
        groupsKernel<<<blocksFor(n, kBlock), kBlock, 0, stream>>>(code, flag.as<uint32_t>(), before.as<uint32_t>(), n,
                                                                  r, group[cur].as<uint32_t>(),
                                                                  starts[cur].as<uint32_t>(),
                                                                  out.cells.as<uint32_t>() + level.first);
        if (r == finest) {
            leafMomentsKernel<<<blocksFor(groups, 64), 64, 0, stream>>>(
                out.positions.as<float4>(), out.shape.as<uint4>(), keep > 0 ? out.sh.as<uint32_t>() : nullptr, n,
                starts[cur].as<uint32_t>(), groups, keep, shWords, stride, moments[cur].as<float>());
        } else {
            const LodLevel& fine = out.levels[r + 1 - coarsest];
            mergeMomentsKernel<<<blocksFor(groups, 64), 64, 0, stream>>>(
                starts[cur].as<uint32_t>(), groups, group[1 - cur].as<uint32_t>(), fine.count,
                moments[1 - cur].as<float>(), stride, moments[cur].as<float>());
        }
        finalizeKernel<<<blocksFor(groups, 64), 64, 0, stream>>>(
            moments[cur].as<float>(), groups, keep, shWords, stride, out.mergedPositions.as<float4>() + level.first,
            out.mergedShape.as<uint4>() + level.first,
            keep > 0 ? out.mergedSh.as<uint32_t>() + uint64_t{level.first} * shWords : nullptr);
        if (r == finest) {
            if (!(s = out.groups.reserve(uint64_t{n} * 4, "lod.groups")).ok()) return s;
            if (!(s = out.starts.reserve(uint64_t{groups} * 4, "lod.finestStarts")).ok()) return s;
            (void)hipMemcpyAsync(out.groups.get(), group[cur].get(), uint64_t{n} * 4, hipMemcpyDeviceToDevice, stream);
            (void)hipMemcpyAsync(out.starts.get(), starts[cur].get(), uint64_t{groups} * 4, hipMemcpyDeviceToDevice,
                                 stream);
        }
        cur = 1 - cur;
    }
    if (!(s = check(hipMemcpyAsync(hostCnt, cnt, sizeof hostCnt, hipMemcpyDeviceToHost, stream), "build")).ok())
        return s;
    if (!(s = check(hipStreamSynchronize(stream), "build")).ok()) return s;
    if (!(s = check(hipGetLastError(), "build")).ok()) return s;
    if (hostCnt[19] != 0) return Status::error(ErrorCode::InternalError, "the scan's look-back gave up");
    return {};
}

struct CutSelector::Frame {
    DeviceBuffer selected, dest, scanStatus, counters;
    DeviceBuffer positions, shape, sh;
    uint32_t elements = 0, capacity = 0, shWords = 0;
    SplatCloud cloud;
};

CutSelector::CutSelector() = default;

CutSelector::~CutSelector() {
    if (hostCounts_ != nullptr) (void)hipHostFree(hostCounts_);
}

Status CutSelector::select(int device, hipStream_t stream, const Projection& projection,
                           std::span<const LodInstance> instances, float threshold, std::vector<SplatInstance>& out,
                           std::vector<CutStats>* stats) {
    out.clear();
    if (stats != nullptr) stats->clear();
    Status s = check(hipSetDevice(device), "hipSetDevice");
    if (!s.ok()) return s;
    if (hostCounts_ == nullptr && !(s = check(hipHostMalloc(reinterpret_cast<void**>(&hostCounts_), 16),
                                             "pinned counts")).ok()) {
        return s;
    }
    while (frames_.size() < instances.size()) frames_.push_back(std::make_unique<Frame>());
    for (size_t k = 0; k < instances.size(); ++k) {
        const LodInstance& instance = instances[k];
        if (instance.cloud == nullptr || instance.cloud->count == 0) {
            if (stats != nullptr) stats->emplace_back();
            continue;
        }
        const LodCloud& lod = *instance.cloud;
        if (lod.levels.empty()) return Status::error(ErrorCode::InvalidArgument, "a cut needs a merged level");
        if (lod.levels.size() > kLevels) return Status::error(ErrorCode::InvalidArgument, "too many levels");
        Frame& frame = *frames_[k];
        const uint32_t total = lod.merged + lod.count;
        if (total > frame.elements) {
            if (!(s = frame.selected.reserve(uint64_t{total} * 4, "cut.selected")).ok()) return s;
            if (!(s = frame.dest.reserve(uint64_t{total} * 4, "cut.dest")).ok()) return s;
            if (!(s = frame.scanStatus.reserve(kern::scanStatusWords(total) * 8, "cut.scanStatus")).ok()) return s;
            if (!(s = frame.counters.reserve(16, "cut.counters")).ok()) return s;
            frame.elements = total;
        }
        // The eye in the cloud's space; perspective size is edge over
        // distance, which a uniform scale leaves alone, orthographic is not
        // (Lod.cpp:410-418).
        const Vec3 eye = transformPoint(inverseAffine(instance.objectToWorld), projection.eyeWorld);
        double scale = 0.0;
        for (int c = 0; c < 3; ++c) {
            const double x = instance.objectToWorld.at(0, c), y = instance.objectToWorld.at(1, c),
                         z = instance.objectToWorld.at(2, c);
            scale = std::max(scale, std::sqrt(x * x + y * y + z * z));
        }
        CutParams p{};
        p.merged = lod.merged;
        p.splats = lod.count;
        p.levels = static_cast<uint32_t>(lod.levels.size());
        p.finestFirst = lod.levels.back().first;
        p.finestLevel = lod.levels.back().level;
        for (size_t l = 0; l < lod.levels.size(); ++l) {
            p.levelOf[l] = lod.levels[l].level;
            p.first[l] = lod.levels[l].first;
        }
        p.first[lod.levels.size()] = lod.merged;
        for (int c = 0; c < 3; ++c) p.lo[c] = lod.boundsLo[c];
        p.extent = lod.extent;
        p.eye[0] = static_cast<float>(eye.x);
        p.eye[1] = static_cast<float>(eye.y);
        p.eye[2] = static_cast<float>(eye.z);
        p.pixelsPerUnit = static_cast<float>(projection.orthographic ? projection.focalX * scale : projection.focalX);
        p.threshold = instance.threshold.value_or(threshold);
        p.orthographic = projection.orthographic ? 1u : 0u;

        uint32_t* cnt = frame.counters.as<uint32_t>();
        kern::fill32(cnt, 0u, 4, stream);   // the scan only ever sets its error word
        cutKernel<<<blocksFor(total, kBlock), kBlock, 0, stream>>>(p, lod.cells.as<uint32_t>(),
                                                                   lod.groups.as<uint32_t>(),
                                                                   frame.selected.as<uint32_t>());
        kern::scanExclusive(frame.selected.as<uint32_t>(), frame.dest.as<uint32_t>(), total,
                            frame.scanStatus.as<uint64_t>(), cnt + 0, cnt + 1, cnt + 2, stream);
        // drawn, merged drawn (the exclusive sum at the first splat), scan error
        (void)hipMemcpyAsync(hostCounts_, cnt, 4, hipMemcpyDeviceToHost, stream);
        (void)hipMemcpyAsync(hostCounts_ + 1, frame.dest.as<uint32_t>() + lod.merged, 4, hipMemcpyDeviceToHost, stream);
        (void)hipMemcpyAsync(hostCounts_ + 2, cnt + 2, 4, hipMemcpyDeviceToHost, stream);
        if (!(s = check(hipStreamSynchronize(stream), "cut")).ok()) return s;
        if (hostCounts_[2] != 0) return Status::error(ErrorCode::InternalError, "the scan's look-back gave up");
        const uint32_t drawn = hostCounts_[0], mergedDrawn = hostCounts_[1];
        if (drawn > frame.capacity || frame.shWords != lod.shWords) {
            const uint32_t capacity = std::max(drawn, frame.capacity + frame.capacity / 2);
            if (!(s = frame.positions.reserve(uint64_t{std::max(capacity, 1u)} * 16, "cut.positions")).ok()) return s;
            if (!(s = frame.shape.reserve(uint64_t{std::max(capacity, 1u)} * 16, "cut.shape")).ok()) return s;
            if (!(s = frame.sh.reserve(uint64_t{std::max(capacity, 1u)} * lod.shWords * 4, "cut.sh")).ok()) return s;
            frame.capacity = std::max(capacity, 1u);
            frame.shWords = lod.shWords;
        }
        if (drawn > 0) {
            gatherKernel<<<blocksFor(total, kBlock), kBlock, 0, stream>>>(
                lod.merged, total, frame.selected.as<uint32_t>(), frame.dest.as<uint32_t>(),
                lod.mergedPositions.as<float4>(), lod.mergedShape.as<uint4>(),
                lod.restPerColour > 0 ? lod.mergedSh.as<uint32_t>() : nullptr, lod.positions.as<float4>(),
                lod.shape.as<uint4>(), lod.restPerColour > 0 ? lod.sh.as<uint32_t>() : nullptr, lod.shWords,
                frame.positions.as<float4>(), frame.shape.as<uint4>(),
                lod.restPerColour > 0 ? frame.sh.as<uint32_t>() : nullptr);
            if (!(s = check(hipGetLastError(), "cut gather")).ok()) return s;
        }
        frame.cloud = SplatCloud{};
        frame.cloud.positions = frame.positions.as<float4>();
        frame.cloud.shape = frame.shape.as<uint4>();
        frame.cloud.sh = lod.restPerColour > 0 ? frame.sh.as<uint32_t>() : nullptr;
        frame.cloud.count = drawn;
        frame.cloud.restPerColour = lod.restPerColour;
        frame.cloud.shWords = lod.shWords;
        if (stats != nullptr) stats->push_back(CutStats{drawn - mergedDrawn, mergedDrawn, lod.count});
        if (drawn > 0) {
            SplatInstance inst;
            inst.splats = &frame.cloud;
            inst.objectToWorld = instance.objectToWorld;
            inst.edit = instance.edit;
            out.push_back(inst);
        }
    }
    return {};
}

}   // namespace lrt::lod
