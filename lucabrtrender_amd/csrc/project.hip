// Preprocess: one thread per splat. EWA projection, the optional Mip-Splatting
// 3D filter and the 2D dilation with its opacity compensation, conic, cutoff,
// tile rectangle, the exact count of tiles the footprint touches, and the
// colour from the fp16 harmonics. Follows the reference's splatProject
// (shaders/lrt/splat/splat_project.slang:52-221; relighting, which is off in
// every frame this path draws, is not carried).
//
// Bandwidth: reads 32 B of packed splat (+ 4 B filter, + 4 * shWords B of
// harmonics at degree > 0) and writes 48 + 8 + 4 B per splat. A workgroup
// first stages its 256 splats' harmonic words through LDS with coalesced
// loads, so the per-thread 92-byte rows cost whole cache lines once.
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "splat_math.h"

namespace lrt::kern {
namespace {

using namespace lrt::dev;

constexpr uint32_t kBlock = 256;
constexpr uint32_t kMaxShWords = 23;   // 45 halves at degree 3, rgb per basis

__global__ __launch_bounds__(kBlock) void projectKernel(FrameParams f, const float4* __restrict__ positions,
                                                        const uint4* __restrict__ shape,
                                                        const uint32_t* __restrict__ sh,
                                                        const float* __restrict__ filter3d,
                                                        float4* __restrict__ proj, uint2* __restrict__ rects,
                                                        uint32_t* __restrict__ touchedOut) {
    __shared__ uint32_t sSh[kBlock * kMaxShWords];
    const uint32_t tid = threadIdx.x;
    const uint32_t first = blockIdx.x * kBlock;
    const uint32_t i = first + tid;
    const uint32_t degree = degreeOf(f.restPerColour) < f.shLimit ? degreeOf(f.restPerColour) : f.shLimit;
    const uint32_t words = degree == 0 ? 0u : f.shWords;
    if (words > 0) {
        // Coalesced copy of this workgroup's harmonic rows into LDS.
        const uint32_t inBlock = f.count - first < kBlock ? f.count - first : kBlock;
        const uint32_t total = inBlock * words;
        const uint32_t* src = sh + static_cast<uint64_t>(first) * words;
        for (uint32_t w = tid; w < total; w += kBlock) sSh[w] = src[w];
    }
    __syncthreads();
    if (i >= f.count) return;
    const uint32_t slot = f.base + i;
    const float4 p = positions[i];
    const uint4 row = shape[i];
    float sx = expF(lowHalf(row.y));
    float sy = expF(highHalf(row.y));
    float sz = expF(lowHalf(row.z));
    const EditFate fate = editFate(f.edit, p.x, p.y, p.z, p.w, sx, sy, sz);
    const uint2 emptyRect = make_uint2(1u | (1u << 16), 0u);
    if (fate.removed) {
        touchedOut[slot] = 0;
        rects[slot] = emptyRect;
        return;
    }
    float opacity = fate.opacity;
    if (f.useFilter3d) {
        // Mip-Splatting: scale^2 + f^2 per axis, opacity * sqrt(det before / det after).
        const float fl = filter3d[i];
        const float f2 = fl * fl;
        const float s2x = sx * sx, s2y = sy * sy, s2z = sz * sz;
        const float a2x = s2x + f2, a2y = s2y + f2, a2z = s2z + f2;
        const float det1 = (s2x * s2y) * s2z;
        const float det2 = (a2x * a2y) * a2z;
        opacity = opacity * __builtin_sqrtf(det1 / det2);
        sx = __builtin_sqrtf(a2x);
        sy = __builtin_sqrtf(a2y);
        sz = __builtin_sqrtf(a2z);
    }
    const float* m = f.m;
    const float tx = m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3];
    const float ty = m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7];
    const float tz = m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11];
    if (!(tz > f.nearZ) || tz > f.farZ) {
        touchedOut[slot] = 0;
        rects[slot] = emptyRect;
        return;
    }
    float q[4];
    decodeQuaternion(row.x, q);
    const float x = q[0], y = q[1], z = q[2], w = q[3];
    const float r00 = 1.0f - 2.0f * (y * y + z * z), r01 = 2.0f * (x * y - w * z), r02 = 2.0f * (x * z + w * y);
    const float r10 = 2.0f * (x * y + w * z), r11 = 1.0f - 2.0f * (x * x + z * z), r12 = 2.0f * (y * z - w * x);
    const float r20 = 2.0f * (x * z - w * y), r21 = 2.0f * (y * z + w * x), r22 = 1.0f - 2.0f * (x * x + y * y);
    // M = R diag(s)
    const float m00 = r00 * sx, m01 = r01 * sy, m02 = r02 * sz;
    const float m10 = r10 * sx, m11 = r11 * sy, m12 = r12 * sz;
    const float m20 = r20 * sx, m21 = r21 * sy, m22 = r22 * sz;

    float jx0, jx2, jy1, jy2, cx, cy;
    if (f.orthographic) {
        jx0 = f.focalX; jx2 = 0.0f;
        jy1 = f.focalY; jy2 = 0.0f;
        cx = f.focalX * tx + f.centreX;
        cy = f.focalY * ty + f.centreY;
    } else {
        const float invZ = 1.0f / tz;
        cx = f.focalX * tx * invZ + f.centreX;
        cy = f.focalY * ty * invZ + f.centreY;
        // The Jacobian is linearised at a point clamped to 1.3x the frustum.
        const float limLoX = (-0.15f * static_cast<float>(f.width) - f.centreX) / f.focalX;
        const float limHiX = (1.15f * static_cast<float>(f.width) - f.centreX) / f.focalX;
        const float limLoY = (-0.15f * static_cast<float>(f.height) - f.centreY) / f.focalY;
        const float limHiY = (1.15f * static_cast<float>(f.height) - f.centreY) / f.focalY;
        const float ctx = clampS(tx * invZ, limLoX, limHiX) * tz;
        const float cty = clampS(ty * invZ, limLoY, limHiY) * tz;
        jx0 = f.focalX * invZ;
        jx2 = -f.focalX * ctx * invZ * invZ;
        jy1 = f.focalY * invZ;
        jy2 = -f.focalY * cty * invZ * invZ;
    }
    // Row vectors J W (the zero entries of J kept, so the sums round as the
    // reference's three-term dot products do), then times M.
    const float t3x0 = jx0 * m[0] + 0.0f * m[4] + jx2 * m[8];
    const float t3x1 = jx0 * m[1] + 0.0f * m[5] + jx2 * m[9];
    const float t3x2 = jx0 * m[2] + 0.0f * m[6] + jx2 * m[10];
    const float t3y0 = 0.0f * m[0] + jy1 * m[4] + jy2 * m[8];
    const float t3y1 = 0.0f * m[1] + jy1 * m[5] + jy2 * m[9];
    const float t3y2 = 0.0f * m[2] + jy1 * m[6] + jy2 * m[10];
    const float ax0 = t3x0 * m00 + t3x1 * m10 + t3x2 * m20;
    const float ax1 = t3x0 * m01 + t3x1 * m11 + t3x2 * m21;
    const float ax2 = t3x0 * m02 + t3x1 * m12 + t3x2 * m22;
    const float ay0 = t3y0 * m00 + t3y1 * m10 + t3y2 * m20;
    const float ay1 = t3y0 * m01 + t3y1 * m11 + t3y2 * m21;
    const float ay2 = t3y0 * m02 + t3y1 * m12 + t3y2 * m22;
    float a = ax0 * ax0 + ax1 * ax1 + ax2 * ax2;
    const float b = ax0 * ay0 + ax1 * ay1 + ax2 * ay2;
    float c = ay0 * ay0 + ay1 * ay1 + ay2 * ay2;

    const float detBefore = a * c - b * b;
    a = a + f.dilation;
    c = c + f.dilation;
    const float det = a * c - b * b;
    if (!(det > 1e-12f)) {
        touchedOut[slot] = 0;
        rects[slot] = emptyRect;
        return;
    }
    float alpha = opacity;
    if (f.antialias) alpha = alpha * __builtin_sqrtf(fmaxS(detBefore, 0.0f) / det);
    if (!(alpha >= 1.0f / 255.0f)) {
        touchedOut[slot] = 0;
        rects[slot] = emptyRect;
        return;
    }
    const float ca = c / det, cb = -b / det, cc = a / det;
    const float cutoff = cutoffPower(alpha);
    uint32_t rect[4];
    tileRect(f.width, f.height, cx, cy, ca, cb, cc, cutoff, rect);
    if (rect[0] > rect[2] || rect[1] > rect[3]) {
        touchedOut[slot] = 0;
        rects[slot] = emptyRect;
        return;
    }
    uint32_t touched = 0;
    for (uint32_t tyy = rect[1]; tyy <= rect[3]; ++tyy)
        for (uint32_t txx = rect[0]; txx <= rect[2]; ++txx)
            touched += tileTouched(f.width, f.height, cx, cy, ca, cb, cc, cutoff, txx, tyy) ? 1u : 0u;
    if (touched == 0) {
        touchedOut[slot] = 0;
        rects[slot] = emptyRect;
        return;
    }

    // Colour from this eye, in the cloud's own space.
    const float dxo = p.x - f.eye[0], dyo = p.y - f.eye[1], dzo = p.z - f.eye[2];
    const float len = __builtin_sqrtf((dxo * dxo + dyo * dyo) + dzo * dzo);
    const float dx = dxo / len, dy = dyo / len, dz = dzo / len;
    // Unpacked with constant indices so the 45 coefficients stay in registers.
    float coeff[45];
    const uint32_t halves = degree == 0 ? 0u : (degree == 1 ? 9u : degree == 2 ? 24u : 45u);
    const uint32_t* shRow = sSh + tid * words;
#pragma unroll
    for (uint32_t k = 0; k < 45; ++k) {
        const uint32_t wd = k < halves ? shRow[k >> 1] : 0u;
        coeff[k] = (k & 1u) == 0u ? lowHalf(wd) : highHalf(wd);
    }
    float rest[3];
    evaluateRest(coeff, degree, dx, dy, dz, rest);
    float colour[3] = {fmaxS(highHalf(row.z) + rest[0], 0.0f), fmaxS(lowHalf(row.w) + rest[1], 0.0f),
                       fmaxS(highHalf(row.w) + rest[2], 0.0f)};
    if (fate.graded) editGrade(f.edit, colour);

    float4* out = proj + static_cast<uint64_t>(slot) * 3;
    out[0] = make_float4(cx, cy, tz, cutoff);
    out[1] = make_float4(ca, cb, cc, alpha);
    out[2] = make_float4(colour[0], colour[1], colour[2], 0.0f);
    rects[slot] = make_uint2(rect[0] | (rect[1] << 16), rect[2] | (rect[3] << 16));
    touchedOut[slot] = touched;
}

}   // namespace

void project(const FrameParams& f, const float4* positions, const uint4* shape, const uint32_t* sh,
             const float* filter3d, float4* proj, uint2* rects, uint32_t* touched, hipStream_t stream) {
    if (f.count == 0) return;
    const uint32_t blocks = (f.count + kBlock - 1) / kBlock;
    projectKernel<<<blocks, kBlock, 0, stream>>>(f, positions, shape, sh, filter3d, proj, rects, touched);
}

}   // namespace lrt::kern
