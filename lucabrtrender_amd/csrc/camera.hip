// Host camera math in FP64, and the per-instance kernel constant block.
// Follows aofx's Transform.h (modules/aofx/sdk/include/aofx/Transform.h:87-419),
// modules/render/src/Camera.cpp:9-48 and FrameParams.h:18-69, in the same
// operation order, so the floats handed to the kernels are the reference's.
#include <cmath>

#include "rasterizer.h"

namespace lrt {
namespace {

constexpr double kToRadians = 3.14159265358979323846 / 180.0;

double length(const Vec3& a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
Vec3 normalised(const Vec3& a) {
    const double l = length(a);
    if (!(l > 1e-12)) return {0.0, 0.0, 0.0};
    const double s = 1.0 / l;
    return {a.x * s, a.y * s, a.z * s};
}
Vec3 cross(const Vec3& a, const Vec3& b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

}   // namespace

Mat4 operator*(const Mat4& a, const Mat4& b) {
    Mat4 out;
    for (int r = 0; r < 4; ++r) {
        for (int c = 0; c < 4; ++c) {
            double sum = 0.0;
            for (int k = 0; k < 4; ++k) sum += a.at(r, k) * b.at(k, c);
            out.at(r, c) = sum;
        }
    }
    return out;
}

Mat4 scaling(const Vec3& s) {
    Mat4 out;
    out.at(0, 0) = s.x;
    out.at(1, 1) = s.y;
    out.at(2, 2) = s.z;
    return out;
}

Mat4 rotationZ(double degrees) {
    const double c = std::cos(degrees * kToRadians);
    const double s = std::sin(degrees * kToRadians);
    Mat4 out;
    out.at(0, 0) = c;
    out.at(0, 1) = -s;
    out.at(1, 0) = s;
    out.at(1, 1) = c;
    return out;
}

Vec3 transformPoint(const Mat4& m, const Vec3& p) {
    return {m.at(0, 0) * p.x + m.at(0, 1) * p.y + m.at(0, 2) * p.z + m.at(0, 3),
            m.at(1, 0) * p.x + m.at(1, 1) * p.y + m.at(1, 2) * p.z + m.at(1, 3),
            m.at(2, 0) * p.x + m.at(2, 1) * p.y + m.at(2, 2) * p.z + m.at(2, 3)};
}

// By cofactors: any 3x3, scale and skew included.
Mat4 inverseAffine(const Mat4& m) {
    const double a00 = m.at(0, 0), a01 = m.at(0, 1), a02 = m.at(0, 2);
    const double a10 = m.at(1, 0), a11 = m.at(1, 1), a12 = m.at(1, 2);
    const double a20 = m.at(2, 0), a21 = m.at(2, 1), a22 = m.at(2, 2);
    const double c00 = a11 * a22 - a12 * a21;
    const double c01 = a12 * a20 - a10 * a22;
    const double c02 = a10 * a21 - a11 * a20;
    const double det = a00 * c00 + a01 * c01 + a02 * c02;
    const double inv = std::fabs(det) > 1e-18 ? 1.0 / det : 0.0;
    Mat4 out;
    out.at(0, 0) = c00 * inv;
    out.at(0, 1) = (a02 * a21 - a01 * a22) * inv;
    out.at(0, 2) = (a01 * a12 - a02 * a11) * inv;
    out.at(1, 0) = c01 * inv;
    out.at(1, 1) = (a00 * a22 - a02 * a20) * inv;
    out.at(1, 2) = (a02 * a10 - a00 * a12) * inv;
    out.at(2, 0) = c02 * inv;
    out.at(2, 1) = (a01 * a20 - a00 * a21) * inv;
    out.at(2, 2) = (a00 * a11 - a01 * a10) * inv;
    const double tx = m.at(0, 3), ty = m.at(1, 3), tz = m.at(2, 3);
    for (int r = 0; r < 3; ++r) out.at(r, 3) = (out.at(r, 0) * tx + out.at(r, 1) * ty + out.at(r, 2) * tz) * -1.0;
    return out;
}

// A -Z look axis: the camera convention.
Mat4 lookAt(const Vec3& eye, const Vec3& target, const Vec3& up) {
    const Vec3 f = normalised({target.x - eye.x, target.y - eye.y, target.z - eye.z});
    Mat4 out;
    if (length(f) < 1e-12) {
        out.at(0, 3) = eye.x;
        out.at(1, 3) = eye.y;
        out.at(2, 3) = eye.z;
        return out;
    }
    const Vec3 zAxis{f.x * -1.0, f.y * -1.0, f.z * -1.0};
    Vec3 xAxis = normalised(cross(up, zAxis));
    if (length(xAxis) < 1e-12) {
        xAxis = normalised(cross(Vec3{0.0, 0.0, 1.0}, zAxis));
        if (length(xAxis) < 1e-12) xAxis = {1.0, 0.0, 0.0};
    }
    const Vec3 yAxis = cross(zAxis, xAxis);
    out.at(0, 0) = xAxis.x; out.at(0, 1) = yAxis.x; out.at(0, 2) = zAxis.x; out.at(0, 3) = eye.x;
    out.at(1, 0) = xAxis.y; out.at(1, 1) = yAxis.y; out.at(1, 2) = zAxis.y; out.at(1, 3) = eye.y;
    out.at(2, 0) = xAxis.z; out.at(2, 1) = yAxis.z; out.at(2, 2) = zAxis.z; out.at(2, 3) = eye.z;
    return out;
}

Camera Camera::lookingAt(const Vec3& eye, const Vec3& target, const Vec3& up) {
    Camera camera;
    camera.cameraToWorld = lookAt(eye, target, up);
    return camera;
}

Projection projectionFor(const Camera& camera, uint32_t width, uint32_t height) {
    Projection out;
    out.worldToView = scaling(Vec3{1.0, 1.0, -1.0}) * inverseAffine(camera.cameraToWorld);
    if (std::fabs(camera.lens.windowRoll) > 1e-12) out.worldToView = rotationZ(camera.lens.windowRoll) * out.worldToView;
    const double w = static_cast<double>(width);
    const double h = static_cast<double>(height);
    const double focal = w * std::fmax(camera.lens.focal, 1e-9) / std::fmax(camera.lens.haperture, 1e-9);
    out.focalX = focal * camera.lens.windowScale[0];
    out.focalY = focal * camera.lens.windowScale[1];
    out.centreX = w * 0.5 + camera.lens.windowTranslate[0] * w * 0.5;
    out.centreY = h * 0.5 + camera.lens.windowTranslate[1] * w * 0.5;
    out.nearZ = camera.lens.nearZ;
    out.farZ = camera.lens.farZ;
    out.orthographic = camera.lens.orthographic;
    out.eyeWorld = {camera.cameraToWorld.at(0, 3), camera.cameraToWorld.at(1, 3), camera.cameraToWorld.at(2, 3)};
    return out;
}

uint32_t bitsFor(uint32_t values) {
    uint32_t bits = 1;
    while (bits < 32 && (uint64_t{1} << bits) < values) ++bits;
    return bits;
}

FrameParams frameParamsFor(const Projection& p, const SplatInstance& instance, const RenderSettings& s,
                           uint32_t base) {
    FrameParams f{};
    f.width = s.width;
    f.height = s.height;
    f.tileShift = tileShiftOf(s.tileSize);
    f.tilesX = tilesAlong(s.width, f.tileShift);
    f.tilesY = tilesAlong(s.height, f.tileShift);
    f.focalX = static_cast<float>(p.focalX);
    f.focalY = static_cast<float>(p.focalY);
    f.centreX = static_cast<float>(p.centreX);
    f.centreY = static_cast<float>(p.centreY);
    f.nearZ = static_cast<float>(p.nearZ);
    f.farZ = static_cast<float>(p.farZ);
    f.limLoX = (-0.15f * static_cast<float>(f.width) - f.centreX) / f.focalX;
    f.limHiX = (1.15f * static_cast<float>(f.width) - f.centreX) / f.focalX;
    f.limLoY = (-0.15f * static_cast<float>(f.height) - f.centreY) / f.focalY;
    f.limHiY = (1.15f * static_cast<float>(f.height) - f.centreY) / f.focalY;
    f.orthographic = p.orthographic ? 1u : 0u;
    f.antialias = s.antialias ? 1u : 0u;
    const Mat4 objectToView = p.worldToView * instance.objectToWorld;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 4; ++c) f.m[r * 4 + c] = static_cast<float>(objectToView.at(r, c));
    const Vec3 eye = transformPoint(inverseAffine(instance.objectToWorld), p.eyeWorld);
    f.eye[0] = static_cast<float>(eye.x);
    f.eye[1] = static_cast<float>(eye.y);
    f.eye[2] = static_cast<float>(eye.z);
    const SplatCloud* cloud = instance.splats;
    f.count = cloud != nullptr ? cloud->count : 0u;
    f.base = base;
    f.restPerColour = cloud != nullptr ? cloud->restPerColour : 0u;
    f.shWords = cloud != nullptr ? cloud->shWords : 1u;
    f.shLimit = s.maxShDegree;
    f.tileBits = bitsFor(f.tilesX * f.tilesY);
    f.depthMode = s.depthMode;
    f.depthThreshold = s.depthThreshold;
    for (int k = 0; k < 4; ++k) f.bg[k] = s.background[static_cast<size_t>(k)];
    f.linearise = s.linearise ? 1u : 0u;
    f.dilation = s.dilation;
    f.useFilter3d = s.filter3d ? 1u : 0u;
    f.edit = instance.edit;
    return f;
}

}   // namespace lrt
