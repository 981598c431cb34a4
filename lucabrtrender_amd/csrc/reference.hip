// The brute-force ground truth on the GPU, in FP64: the reference's
// ReferenceRenderer (modules/render/src/ReferenceRenderer.cpp:54-160;
// shaders/lrt/reference/reference_project.slang:31-105, reference_blend.slang:
// 129-223) without its per-pixel hit cap. Every splat is projected in double;
// the visible ones are sorted by (view z, index) with the library's radix sort
// on the doubles' bits (positive doubles order as their bits), or by the
// rasteriser's own 24-bit depth key; then each pixel walks all of them in that
// order with the reference's skip and stop rules, in double.
//
// This is the error oracle for whole frames: images of the tile rasteriser
// are compared with it at full size. It is O(pixels x splats) by design.
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "splat_math.h"

namespace lrt::kern {
namespace {

using namespace lrt::dev;

constexpr uint32_t kBlock = 256;

struct RefRecord {   // 12 doubles
    double cx, cy, z, cutoff;
    double a, b, c, alpha;
    double r, g, bl, pad;
};

__device__ __forceinline__ double halfD(uint32_t w, bool high) {
    return static_cast<double>(high ? highHalf(w) : lowHalf(w));
}

__global__ __launch_bounds__(kBlock) void refProjectKernel(FrameParams f, const float4* __restrict__ positions,
                                                           const uint4* __restrict__ shape,
                                                           const uint32_t* __restrict__ sh,
                                                           const float* __restrict__ filter3d,
                                                           RefRecord* __restrict__ out, uint64_t* __restrict__ keys,
                                                           uint32_t* __restrict__ values, uint32_t* __restrict__ live,
                                                           int keyed) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= f.count) return;
    const uint32_t slot = f.base + i;
    keys[slot] = ~0ull;   // culled splats sort last
    values[slot] = slot;
    const float4 pf = positions[i];
    const double p[3] = {pf.x, pf.y, pf.z};
    const uint4 row = shape[i];
    double s[3] = {exp(halfD(row.y, false)), exp(halfD(row.y, true)), exp(halfD(row.z, false))};
    const EditFate fate = editFate(f.edit, pf.x, pf.y, pf.z, pf.w, static_cast<float>(s[0]),
                                   static_cast<float>(s[1]), static_cast<float>(s[2]));
    if (fate.removed) return;
    double opacity = fate.graded ? static_cast<double>(pf.w) * static_cast<double>(f.edit.opacity)
                                 : static_cast<double>(pf.w);
    if (f.useFilter3d) {
        const double fl = filter3d[i], f2 = fl * fl;
        double det1 = 1.0, det2 = 1.0;
        for (int k = 0; k < 3; ++k) {
            det1 *= s[k] * s[k];
            det2 *= s[k] * s[k] + f2;
            s[k] = sqrt(s[k] * s[k] + f2);
        }
        opacity *= sqrt(det1 / det2);
    }
    double m[12];
    for (int k = 0; k < 12; ++k) m[k] = f.m[k];
    const double v[3] = {m[0] * p[0] + m[1] * p[1] + m[2] * p[2] + m[3],
                         m[4] * p[0] + m[5] * p[1] + m[6] * p[2] + m[7],
                         m[8] * p[0] + m[9] * p[1] + m[10] * p[2] + m[11]};
    if (!(v[2] > static_cast<double>(f.nearZ)) || v[2] > static_cast<double>(f.farZ)) return;
    float qf[4];
    decodeQuaternion(row.x, qf);
    double q[4] = {qf[0], qf[1], qf[2], qf[3]};
    const double ql = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    for (int k = 0; k < 4; ++k) q[k] /= ql;
    const double x = q[0], y = q[1], z = q[2], w = q[3];
    const double r[3][3] = {
        {1.0 - 2.0 * (y * y + z * z), 2.0 * (x * y - w * z), 2.0 * (x * z + w * y)},
        {2.0 * (x * y + w * z), 1.0 - 2.0 * (x * x + z * z), 2.0 * (y * z - w * x)},
        {2.0 * (x * z - w * y), 2.0 * (y * z + w * x), 1.0 - 2.0 * (x * x + y * y)}};
    double rs[3][3], cov3[3][3], tmp[3][3], cv[3][3];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) rs[a][b] = r[a][b] * s[b];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) cov3[a][b] = rs[a][0] * rs[b][0] + rs[a][1] * rs[b][1] + rs[a][2] * rs[b][2];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
            tmp[a][b] = cov3[a][0] * m[b * 4 + 0] + cov3[a][1] * m[b * 4 + 1] + cov3[a][2] * m[b * 4 + 2];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
            cv[a][b] = m[a * 4 + 0] * tmp[0][b] + m[a * 4 + 1] * tmp[1][b] + m[a * 4 + 2] * tmp[2][b];
    double j[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
    double cx, cy;
    const double fx = f.focalX, fy = f.focalY, ccx = f.centreX, ccy = f.centreY;
    if (f.orthographic) {
        cx = fx * v[0] + ccx;
        cy = fy * v[1] + ccy;
        j[0][0] = fx;
        j[1][1] = fy;
    } else {
        cx = fx * v[0] / v[2] + ccx;
        cy = fy * v[1] / v[2] + ccy;
        const double lox = (-0.15 * static_cast<double>(f.width) - ccx) / fx;
        const double hix = (1.15 * static_cast<double>(f.width) - ccx) / fx;
        const double loy = (-0.15 * static_cast<double>(f.height) - ccy) / fy;
        const double hiy = (1.15 * static_cast<double>(f.height) - ccy) / fy;
        double tx = v[0] / v[2], ty = v[1] / v[2];
        tx = tx < lox ? lox : tx > hix ? hix : tx;
        ty = ty < loy ? loy : ty > hiy ? hiy : ty;
        j[0][0] = fx / v[2];
        j[0][2] = -fx * (tx * v[2]) / (v[2] * v[2]);
        j[1][1] = fy / v[2];
        j[1][2] = -fy * (ty * v[2]) / (v[2] * v[2]);
    }
    double t2[3][3], c2[3][3];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) t2[a][b] = cv[a][0] * j[b][0] + cv[a][1] * j[b][1] + cv[a][2] * j[b][2];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) c2[a][b] = j[a][0] * t2[0][b] + j[a][1] * t2[1][b] + j[a][2] * t2[2][b];
    const double dil = f.dilation;
    const double before = c2[0][0] * c2[1][1] - c2[0][1] * c2[0][1];
    const double a = c2[0][0] + dil, b = c2[0][1], c = c2[1][1] + dil;
    const double det = a * c - b * b;
    if (!(det > 1e-12)) return;
    double alpha = opacity;
    if (f.antialias) alpha *= sqrt((before > 0.0 ? before : 0.0) / det);
    if (!(alpha >= 1.0 / 255.0)) return;
    const double cut = 2.0 * log(255.0 * alpha > 1.0001 ? 255.0 * alpha : 1.0001);
    const uint32_t degree = degreeOf(f.restPerColour) < f.shLimit ? degreeOf(f.restPerColour) : f.shLimit;
    const double dxo = p[0] - f.eye[0], dyo = p[1] - f.eye[1], dzo = p[2] - f.eye[2];
    const double len = sqrt(dxo * dxo + dyo * dyo + dzo * dzo);
    const double dx = dxo / len, dy = dyo / len, dz = dzo / len;
    const double base[3] = {halfD(row.z, true), halfD(row.w, false), halfD(row.w, true)};
    const uint32_t* shRow = sh + static_cast<uint64_t>(i) * f.shWords;
    double colour[3];
    for (int ch = 0; ch < 3; ++ch) {
        double rest = 0.0;
        const auto bd = [&](uint32_t k) {
            const uint32_t at = k * 3u + static_cast<uint32_t>(ch);
            return halfD(shRow[at >> 1], (at & 1u) != 0u);
        };
        if (degree >= 1) {
            rest = -0.4886025119029199 * dy * bd(0) + 0.4886025119029199 * dz * bd(1) - 0.4886025119029199 * dx * bd(2);
            if (degree >= 2) {
                const double xx = dx * dx, yy = dy * dy, zz = dz * dz;
                rest += 1.0925484305920792 * dx * dy * bd(3) - 1.0925484305920792 * dy * dz * bd(4) +
                        0.31539156525252005 * (2.0 * zz - xx - yy) * bd(5) - 1.0925484305920792 * dx * dz * bd(6) +
                        0.5462742152960396 * (xx - yy) * bd(7);
                if (degree >= 3) {
                    rest += -0.5900435899266435 * dy * (3.0 * xx - yy) * bd(8) +
                            2.890611442640554 * dx * dy * dz * bd(9) -
                            0.4570457994644658 * dy * (4.0 * zz - xx - yy) * bd(10) +
                            0.3731763325901154 * dz * (2.0 * zz - 3.0 * xx - 3.0 * yy) * bd(11) -
                            0.4570457994644658 * dx * (4.0 * zz - xx - yy) * bd(12) +
                            1.445305721320277 * dz * (xx - yy) * bd(13) -
                            0.5900435899266435 * dx * (xx - 3.0 * yy) * bd(14);
                }
            }
        }
        const double v0 = base[ch] + rest;
        colour[ch] = v0 > 0.0 ? v0 : 0.0;
    }
    if (fate.graded) {
        const double tr = colour[0] * f.edit.tint[0] * f.edit.brightness;
        const double tg = colour[1] * f.edit.tint[1] * f.edit.brightness;
        const double tb = colour[2] * f.edit.tint[2] * f.edit.brightness;
        const double luma = tr * 0.2126 + tg * 0.7152 + tb * 0.0722;
        const double t3[3] = {tr, tg, tb};
        for (int k = 0; k < 3; ++k) {
            const double g = luma + (t3[k] - luma) * f.edit.saturation;
            colour[k] = g > 0.0 ? g : 0.0;
        }
    }
    out[slot] = RefRecord{cx, cy, v[2], cut < 9.0 ? cut : 9.0, c / det, -b / det, a / det, alpha,
                          colour[0], colour[1], colour[2], 0.0};
    keys[slot] = keyed ? static_cast<uint64_t>(floatToBits(static_cast<float>(v[2])) >> 7) << 32
                       : static_cast<uint64_t>(__double_as_longlong(v[2]));
    // One add a wave for its live lanes (the ones that reach here).
    const uint64_t liveLanes = __ballot(1);
    if (__lane_id() == static_cast<uint32_t>(__builtin_ctzll(liveLanes))) {
        atomicAdd(live, static_cast<uint32_t>(__popcll(liveLanes)));
    }
}

__device__ __forceinline__ double srgbToLinearF64(double c) {
    c = c > 0.0 ? c : 0.0;
    return c <= 0.04045 ? c / 12.92 : pow((c + 0.055) / 1.055, 2.4);
}

// One thread a pixel: every visible splat in order (the same for the whole
// workgroup, so each record is one uniform load).
__global__ __launch_bounds__(kBlock) void refBlendKernel(FrameParams f, const RefRecord* __restrict__ recs,
                                                         const uint32_t* __restrict__ order,
                                                         const uint32_t* __restrict__ liveCount,
                                                         const uint32_t* __restrict__ pixelList, uint32_t listCount,
                                                         double* __restrict__ colour, double* __restrict__ depth) {
    // Output n is pixel n of the frame, or pixel pixelList[n] when a list is given.
    const uint32_t n = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t pixels = f.width * f.height;
    const uint32_t live = *liveCount;
    const bool inside = pixelList != nullptr ? n < listCount && pixelList[n] < pixels : n < pixels;
    const uint32_t at = pixelList != nullptr ? (n < listCount ? pixelList[n] : 0u) : n;
    const double pxf = static_cast<double>(inside ? at % f.width : 0u) + 0.5;
    const double pyf = static_cast<double>(inside ? at / f.width : 0u) + 0.5;
    double r = 0.0, g = 0.0, b = 0.0, t = 1.0, dsum = 0.0, dat = -1.0;
    bool active = inside;
    for (uint32_t k = 0; k < live; ++k) {
        // Every 64 splats the workgroup leaves once all its pixels are opaque.
        if ((k & 63u) == 0u && __syncthreads_count(active ? 1 : 0) == 0) break;
        if (!active) continue;
        const RefRecord s = recs[order[k]];
        const double dx = pxf - s.cx, dy = pyf - s.cy;
        const double power = s.a * dx * dx + 2.0 * s.b * dx * dy + s.c * dy * dy;
        if (power > s.cutoff) continue;
        double alpha = s.alpha * exp(-0.5 * power);
        alpha = alpha < 0.99 ? alpha : 0.99;
        if (alpha < 1.0 / 255.0) continue;
        const double weight = t * alpha;
        r += weight * s.r;
        g += weight * s.g;
        b += weight * s.bl;
        dsum += weight * s.z;
        t *= 1.0 - alpha;
        if (dat < 0.0 && 1.0 - t >= static_cast<double>(f.depthThreshold)) dat = s.z;
        if (t < 1e-4) active = false;
    }
    if (!inside) return;
    const double coverage = 1.0 - t;
    if (f.linearise && coverage > 1e-6) {
        r = srgbToLinearF64(r / coverage) * coverage;
        g = srgbToLinearF64(g / coverage) * coverage;
        b = srgbToLinearF64(b / coverage) * coverage;
    }
    colour[static_cast<uint64_t>(n) * 4 + 0] = r + t * f.bg[0];
    colour[static_cast<uint64_t>(n) * 4 + 1] = g + t * f.bg[1];
    colour[static_cast<uint64_t>(n) * 4 + 2] = b + t * f.bg[2];
    colour[static_cast<uint64_t>(n) * 4 + 3] = coverage + t * f.bg[3];
    if (depth != nullptr) {
        depth[n] = f.depthMode == 0 ? (coverage > 1e-6 ? dsum / coverage : 0.0) : (dat > 0.0 ? dat : 0.0);
    }
}

}   // namespace

uint64_t referenceRecordBytes() { return sizeof(RefRecord); }

void referenceProject(const FrameParams& f, const float4* positions, const uint4* shape, const uint32_t* sh,
                      const float* filter3d, void* records, uint64_t* keys, uint32_t* values, uint32_t* live,
                      bool keyed, hipStream_t stream) {
    if (f.count == 0) return;
    refProjectKernel<<<(f.count + kBlock - 1) / kBlock, kBlock, 0, stream>>>(
        f, positions, shape, sh, filter3d, static_cast<RefRecord*>(records), keys, values, live, keyed ? 1 : 0);
}

void referenceBlend(const FrameParams& f, const void* records, const uint32_t* order, const uint32_t* live,
                    double* colour, double* depth, hipStream_t stream, const uint32_t* pixelList,
                    uint32_t listCount) {
    const uint32_t pixels = pixelList != nullptr ? listCount : f.width * f.height;
    if (pixels == 0) return;
    refBlendKernel<<<(pixels + kBlock - 1) / kBlock, kBlock, 0, stream>>>(
        f, static_cast<const RefRecord*>(records), order, live, pixelList, listCount, colour, depth);
}

}   // namespace lrt::kern
