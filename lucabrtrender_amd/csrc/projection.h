// The EWA projection of one splat, shared by the packed-cloud preprocess and
// the differentiable (FP32-parameter) one: view position and cull, the
// Jacobian at the point clamped to 1.3x the frustum, the 2D covariance of
// J W R S, the dilation with its opacity compensation, the conic, the cutoff,
// the tile rectangle and the exact count of tiles the footprint touches.
// Follows the reference's splatProject (shaders/lrt/splat/splat_project.slang:
// 52-221) in its own operation order, so the packed path stays bit-exact with
// the CPU restatement.
#pragma once

#include <hip/hip_runtime.h>

#include "frame.h"
#include "splat_math.h"

namespace lrt::dev {

// What one splat brings to the projection, already decoded.
struct SplatShape {
    float px, py, pz;    // centre in the cloud's own space
    float opacity;       // after any edit and the 3D filter
    float sx, sy, sz;    // scales after the 3D filter
    float q[4];          // unit quaternion x, y, z, w
};

// Mip-Splatting's 3D filter: scale^2 + f^2 per axis, opacity * sqrt(det before
// / det after).
__device__ __forceinline__ void applyFilter3d(float fl, float& sx, float& sy, float& sz, float& opacity) {
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

// The rotation matrix of a unit quaternion (x, y, z, w), row major.
__device__ __forceinline__ void rotationOf(const float q[4], float r[9]) {
    const float x = q[0], y = q[1], z = q[2], w = q[3];
    r[0] = 1.0f - 2.0f * (y * y + z * z); r[1] = 2.0f * (x * y - w * z); r[2] = 2.0f * (x * z + w * y);
    r[3] = 2.0f * (x * y + w * z); r[4] = 1.0f - 2.0f * (x * x + z * z); r[5] = 2.0f * (y * z - w * x);
    r[6] = 2.0f * (x * z - w * y); r[7] = 2.0f * (y * z + w * x); r[8] = 1.0f - 2.0f * (x * x + y * y);
}

// What the projection keeps of a drawn splat before its colour.
struct Footprint {
    float cx, cy, tz, cutoff;   // screen centre, view depth, power cutoff
    float ca, cb, cc, alpha;    // conic and opacity
    uint32_t rect[4];           // inclusive tile rectangle x0, y0, x1, y1
};

// The splat's footprint from frame f; false when it is culled or its tile
// rectangle is empty (and then fp is partly written). The tiles it touches are
// counted apart, wave-wide (countTouchedTilesWave).
__device__ __forceinline__ bool footprintOf(const FrameParams& f, const SplatShape& s, Footprint& fp) {
    const float* m = f.m;
    const float tx = m[0] * s.px + m[1] * s.py + m[2] * s.pz + m[3];
    const float ty = m[4] * s.px + m[5] * s.py + m[6] * s.pz + m[7];
    const float tz = m[8] * s.px + m[9] * s.py + m[10] * s.pz + m[11];
    if (!(tz > f.nearZ) || tz > f.farZ) return false;
    fp.tz = tz;
    float r[9];
    rotationOf(s.q, r);
    // M = R diag(s)
    const float m00 = r[0] * s.sx, m01 = r[1] * s.sy, m02 = r[2] * s.sz;
    const float m10 = r[3] * s.sx, m11 = r[4] * s.sy, m12 = r[5] * s.sz;
    const float m20 = r[6] * s.sx, m21 = r[7] * s.sy, m22 = r[8] * s.sz;

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
        const float ctx = clampS(tx * invZ, f.limLoX, f.limHiX) * tz;
        const float cty = clampS(ty * invZ, f.limLoY, f.limHiY) * tz;
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
    if (!(det > 1e-12f)) return false;
    float alpha = s.opacity;
    if (f.antialias) alpha = alpha * __builtin_sqrtf(fmaxS(detBefore, 0.0f) / det);
    if (!(alpha >= 1.0f / 255.0f)) return false;
    const float ca = c / det, cb = -b / det, cc = a / det;
    const float cutoff = cutoffPower(alpha);
    tileRect(f.width, f.height, cx, cy, ca, cb, cc, cutoff, fp.rect, f.tileShift);
    if (fp.rect[0] > fp.rect[2] || fp.rect[1] > fp.rect[3]) return false;
    fp.cx = cx;
    fp.cy = cy;
    fp.cutoff = cutoff;
    fp.ca = ca;
    fp.cb = cb;
    fp.cc = cc;
    fp.alpha = alpha;
    return true;
}

// The unit direction from the eye to the splat's centre, in the cloud's own
// space: the argument of the colour's harmonics.
__device__ __forceinline__ void viewDirection(const FrameParams& f, const SplatShape& s, float& dx, float& dy,
                                              float& dz) {
    const float dxo = s.px - f.eye[0], dyo = s.py - f.eye[1], dzo = s.pz - f.eye[2];
    const float len = __builtin_sqrtf((dxo * dxo + dyo * dyo) + dzo * dzo);
    dx = dxo / len;
    dy = dyo / len;
    dz = dzo / len;
}

// A drawn splat's record but its colour (float4s 0 and 1 of the three) and its
// packed tile rectangle.
__device__ __forceinline__ void writeFootprint(const Footprint& fp, float4* __restrict__ out,
                                               uint2* __restrict__ rectOut) {
    out[0] = make_float4(fp.cx, fp.cy, fp.tz, fp.cutoff);
    out[1] = make_float4(fp.ca, fp.cb, fp.cc, fp.alpha);
    *rectOut = make_uint2(fp.rect[0] | (fp.rect[1] << 16), fp.rect[2] | (fp.rect[3] << 16));
}

// The three float4 of a drawn splat's record and its packed tile rectangle.
__device__ __forceinline__ void writeProjection(const Footprint& fp, const float colour[3], float4* __restrict__ out,
                                                uint2* __restrict__ rectOut) {
    writeFootprint(fp, out, rectOut);
    out[2] = make_float4(colour[0], colour[1], colour[2], 0.0f);
}

// Projects one splat a lane, `want` false for a lane with none; returns the
// tiles its footprint touches (0: culled, and then the record is left alone).
// `colourOf(dx, dy, dz, rgb)` gives the colour for the unit direction from the
// eye, called only for a splat that is drawn; *tz is its view depth. Every
// lane of the wave must reach the call: the tiles are counted wave-wide
// (countTouchedTilesWave).
template <class ColourOf>
__device__ __forceinline__ uint32_t projectWave(const FrameParams& f, const SplatShape& s, bool want,
                                                ColourOf colourOf, float4* __restrict__ out,
                                                uint2* __restrict__ rectOut, float* tz) {
    Footprint fp{0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, {1u, 1u, 0u, 0u}};
    const bool candidate = want && footprintOf(f, s, fp);
    const uint32_t touched = countTouchedTilesWave(f.width, f.height, fp.cx, fp.cy, fp.ca, fp.cb, fp.cc, fp.cutoff,
                                                   fp.rect, candidate, f.tileShift);
    *tz = fp.tz;
    if (touched == 0) return 0;
    float dx, dy, dz;
    viewDirection(f, s, dx, dy, dz);
    float colour[3];
    colourOf(dx, dy, dz, colour);
    writeProjection(fp, colour, out, rectOut);
    return touched;
}

}   // namespace lrt::dev
