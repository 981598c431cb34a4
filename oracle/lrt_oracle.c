/*
 * lrt_oracle -- TEST INFRASTRUCTURE, NOT PRODUCT CODE. See lrt_oracle.h for
 * what this file is, who may load it and what pins each part.
 *
 * Built with -ffp-contract=off: every a*b+c below rounds twice, as the
 * reference's Slang source reads. Evaluation order follows the source
 * expressions left to right.
 */
#include "lrt_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static inline uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static inline float fmin_(float a, float b) { return a < b ? a : b; }
static inline float fmax_(float a, float b) { return a > b ? a : b; }
static inline float clampf_(float v, float lo, float hi) { return fmin_(fmax_(v, lo), hi); }

#ifdef _OPENMP
#define OR_THREADS(t) num_threads((t) > 0 ? (t) : omp_get_max_threads())
#endif

/* ====================================================================== */
/* exp and log. The reference takes both from Slang intrinsics whose bits no */
/* reference test pins; the build defines them here by a published method   */
/* (Cody & Waite argument reduction, Taylor polynomial in Horner form) and  */
/* the HIP library implements the identical sequence of float operations.   */
/* ====================================================================== */

#define OR_LN2_HI 0.693145751953125f          /* ln 2 to 15 significant bits */
#define OR_LN2_LO 1.428606765330187e-06f      /* ln 2 - OR_LN2_HI */
#define OR_LOG2E 1.4426950408889634f

float or_expf(float x) {
    if (x != x) return x;
    if (x > 88.72283935546875f) return INFINITY;
    if (x < -103.97208404541016f) return 0.0f;
    const float kf = rintf(x * OR_LOG2E);
    int k = (int)kf;
    float r = x - kf * OR_LN2_HI;             /* exact: |k| < 2^8, OR_LN2_HI has 15 bits */
    r = r - kf * OR_LN2_LO;
    float p = 1.0f / 5040.0f;
    p = p * r + 1.0f / 720.0f;
    p = p * r + 1.0f / 120.0f;
    p = p * r + 1.0f / 24.0f;
    p = p * r + 1.0f / 6.0f;
    p = p * r + 0.5f;
    p = p * r + 1.0f;
    p = p * r + 1.0f;
    if (k > 127) {
        p = p * 2.0f;
        k -= 1;
    }
    if (k >= -126) {
        return p * u2f((uint32_t)(k + 127) << 23);
    }
    /* Subnormal result: one exact scaling, then the single rounding. */
    return (p * u2f((uint32_t)(k + 64 + 127) << 23)) * u2f((uint32_t)(127 - 64) << 23);
}

float or_logf(float x) {
    if (x != x) return x;
    if (x < 0.0f) return NAN;
    if (x == 0.0f) return -INFINITY;
    uint32_t b = f2u(x);
    if (b == 0x7F800000u) return x;
    int e = 0;
    if (b < 0x00800000u) {
        x = x * 33554432.0f;                  /* 2^25 */
        b = f2u(x);
        e = -25;
    }
    e += (int)(b >> 23) - 127;
    float m = u2f((b & 0x007FFFFFu) | 0x3F800000u);
    if (m > 1.41421353816986083984375f) {
        m = m * 0.5f;
        e += 1;
    }
    /* log(1+f) = 2 atanh(s), s = f/(2+f); 2s = f - s f. */
    const float f = m - 1.0f;
    const float s = f / (2.0f + f);
    const float z = s * s;
    const float r = z * (2.0f / 3.0f + z * (2.0f / 5.0f + z * (2.0f / 7.0f + z * (2.0f / 9.0f + z * (2.0f / 11.0f)))));
    const float lm = f - s * (f - r);
    const float ef = (float)e;
    return ef * OR_LN2_HI + (ef * OR_LN2_LO + lm);
}

static float or_powf(float y, float p) {
    return or_expf(p * or_logf(y));
}

float or_half_to_float(uint16_t h) {
    const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    const uint32_t e = (h >> 10) & 31u;
    const uint32_t m = h & 1023u;
    if (e == 0) {
        const float v = (float)m * 5.9604644775390625e-08f;   /* 2^-24, exact */
        return u2f(f2u(v) | sign);
    }
    if (e == 31) {
        return u2f(sign | 0x7F800000u | (m << 13));
    }
    return u2f(sign | ((e + 112u) << 23) | (m << 13));
}

uint16_t or_float_to_half(float f) {
    const uint32_t x = f2u(f);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const uint32_t a = x & 0x7FFFFFFFu;
    if (a >= 0x7F800000u) {
        return (uint16_t)(sign | 0x7C00u | (a > 0x7F800000u ? (0x200u | ((a >> 13) & 0x3FFu)) : 0u));
    }
    if (a >= 0x477FF000u) {
        return (uint16_t)(sign | 0x7C00u);
    }
    if (a < 0x38800000u) {
        const float v = u2f(a) * 16777216.0f;                  /* in units of 2^-24 */
        return (uint16_t)(sign | (uint32_t)rintf(v));
    }
    uint32_t h = (((a >> 23) - 112u) << 10) | ((a & 0x7FFFFFu) >> 13);
    const uint32_t rest = a & 0x1FFFu;
    if (rest > 0x1000u || (rest == 0x1000u && (h & 1u))) {
        h += 1u;
    }
    return (uint16_t)(sign | h);
}

static inline float low_half(uint32_t w) { return or_half_to_float((uint16_t)(w & 0xFFFFu)); }
static inline float high_half(uint32_t w) { return or_half_to_float((uint16_t)(w >> 16)); }

/* ====================================================================== */
/* packing.slang                                                           */
/* ====================================================================== */

#define OR_SQRT_HALF 0.70710678f
#define OR_SH0 0.28209479177387814f

/* packing.slang:26-50 (rounding to nearest, ties to even) */
uint32_t or_encode_quaternion(const float qin[4]) {
    const float len = sqrtf(((qin[0] * qin[0] + qin[1] * qin[1]) + qin[2] * qin[2]) + qin[3] * qin[3]);
    float q[4] = {qin[0] / len, qin[1] / len, qin[2] / len, qin[3] / len};
    uint32_t largest = 0;
    float best = fabsf(q[0]);
    for (uint32_t k = 1; k < 4; ++k) {
        if (fabsf(q[k]) > best) {
            best = fabsf(q[k]);
            largest = k;
        }
    }
    if (q[largest] < 0.0f) {
        for (int k = 0; k < 4; ++k) q[k] = -q[k];
    }
    uint32_t word = largest;
    uint32_t shift = 22;
    for (uint32_t k = 0; k < 4; ++k) {
        if (k == largest) continue;
        const float unit = clampf_(q[k] / OR_SQRT_HALF * 0.5f + 0.5f, 0.0f, 1.0f);
        word |= (uint32_t)rintf(unit * 1023.0f) << shift;
        shift -= 10;
    }
    return word;
}

/* packing.slang:52-69 */
void or_decode_quaternion(uint32_t word, float q[4]) {
    const uint32_t largest = word & 3u;
    float s0 = (float)((word >> 22) & 1023u);
    float s1 = (float)((word >> 12) & 1023u);
    float s2 = (float)((word >> 2) & 1023u);
    s0 = (s0 / 1023.0f * 2.0f - 1.0f) * OR_SQRT_HALF;
    s1 = (s1 / 1023.0f * 2.0f - 1.0f) * OR_SQRT_HALF;
    s2 = (s2 / 1023.0f * 2.0f - 1.0f) * OR_SQRT_HALF;
    const float rest = sqrtf(fmax_(0.0f, 1.0f - ((s0 * s0 + s1 * s1) + s2 * s2)));
    const float small[3] = {s0, s1, s2};
    uint32_t next = 0;
    for (uint32_t k = 0; k < 4; ++k) {
        if (k == largest) {
            q[k] = rest;
        } else {
            q[k] = small[next++];
        }
    }
}

uint32_t or_pack_halves(float lo, float hi) {
    return (uint32_t)or_float_to_half(lo) | ((uint32_t)or_float_to_half(hi) << 16);
}

static inline float half_safe(float v) { return clampf_(v, -65000.0f, 65000.0f); }

/* splat_encoding.slang:28-96, the modes or_encoding names. */
static float decode_opacity(float v, uint32_t mode) {
    if (mode == 0) return 1.0f / (1.0f + or_expf(-v));
    if (mode == 2) return v / 255.0f;
    return v;
}
static float decode_log_scale(float v, uint32_t mode) {
    if (mode == 2) return v / 16.0f - 10.0f;
    return mode == 0 ? v : or_logf(fmax_(v, 1e-30f));
}
static float decode_colour(float v, uint32_t mode) {
    if (mode == 0) return 0.5f + OR_SH0 * v;
    if (mode == 2) return v / 255.0f;
    if (mode == 3) return 0.5f + OR_SH0 * ((v / 255.0f - 0.5f) / 0.15f);
    return v;
}
static float decode_byte_unit(float v, uint32_t mode) { return mode == 1 ? (v - 128.0f) / 128.0f : v; }

/* splat_encoding.slang restFlipSign: the sign a rest coefficient takes when y
 * and z change sign, per basis in 3DGS order. */
static const float kRestFlip[15] = {-1.0f, -1.0f, 1.0f, -1.0f, 1.0f, 1.0f, -1.0f, 1.0f,
                                    -1.0f, 1.0f, -1.0f, -1.0f, 1.0f, -1.0f, 1.0f};

/* splat_encoding.slang decodeQuaternion: (x, y, z, w), not normalised. */
static void decode_rotation(const float* r, const or_encoding* e, float q[4]) {
    if (e->rotation_mode == 2) {
        q[0] = r[e->rot_x] / 127.5f - 1.0f;
        q[1] = r[e->rot_y] / 127.5f - 1.0f;
        q[2] = r[e->rot_z] / 127.5f - 1.0f;
        q[3] = sqrtf(fmax_(0.0f, 1.0f - ((q[0] * q[0] + q[1] * q[1]) + q[2] * q[2])));
    } else if (e->rotation_mode == 3) {
        uint32_t comp = (uint32_t)r[e->rot_x] | ((uint32_t)r[e->rot_y] << 16);
        const uint32_t largest = comp >> 30;
        float sum = 0.0f;
        for (int i = 3; i >= 0; --i) {
            if ((uint32_t)i == largest) continue;
            const float magnitude = OR_SQRT_HALF * (float)(comp & 511u) / 511.0f;
            q[i] = ((comp >> 9) & 1u) ? -magnitude : magnitude;
            sum = sum + q[i] * q[i];
            comp >>= 10;
        }
        q[largest] = sqrtf(fmax_(0.0f, 1.0f - sum));
    } else {
        q[0] = decode_byte_unit(r[e->rot_x], e->rotation_mode);
        q[1] = decode_byte_unit(r[e->rot_y], e->rotation_mode);
        q[2] = decode_byte_unit(r[e->rot_z], e->rotation_mode);
        q[3] = decode_byte_unit(r[e->rot_w], e->rotation_mode);
    }
    if (e->flip_yz) {
        q[1] = -q[1];
        q[2] = -q[2];
    }
}

/* splat_validate.slang:16-29, splat_decode.slang:17-87: kept when the
 * position and log scales are finite, |log scale| < 60 and opacity >= 1/255;
 * compacted in record order (GpuClouds.cpp:151-185). */
uint32_t or_decode_records(const float* records, uint32_t count, const or_encoding* e,
                           uint32_t keep, float* positions, uint32_t* shape, uint32_t* sh,
                           uint32_t sh_words) {
    uint32_t kept = 0;
    for (uint32_t i = 0; i < count; ++i) {
        const float* r = records + (size_t)i * e->floats_per_record;
        const float scale = 1.0f / (float)(1u << (e->fractional_bits < 24u ? e->fractional_bits : 24u));
        float px = r[e->x] * scale, py = r[e->y] * scale, pz = r[e->z] * scale;
        if (e->flip_yz) {
            py = -py;
            pz = -pz;
        }
        const float a = decode_opacity(r[e->opacity], e->opacity_mode);
        const float l0 = decode_log_scale(r[e->scale0], e->scale_mode);
        const float l1 = decode_log_scale(r[e->scale1], e->scale_mode);
        const float l2 = decode_log_scale(r[e->scale2], e->scale_mode);
        const int finite = isfinite(px) && isfinite(py) && isfinite(pz) && isfinite(l0) && isfinite(l1) &&
                           isfinite(l2) && fabsf(l0) < 60.0f && fabsf(l1) < 60.0f && fabsf(l2) < 60.0f;
        if (!(finite && a >= 1.0f / 255.0f)) continue;
        const uint32_t j = kept++;
        positions[j * 4 + 0] = px;
        positions[j * 4 + 1] = py;
        positions[j * 4 + 2] = pz;
        positions[j * 4 + 3] = clampf_(a, 0.0f, 1.0f);
        float q[4];
        decode_rotation(r, e, q);
        const float qlen = sqrtf(((q[0] * q[0] + q[1] * q[1]) + q[2] * q[2]) + q[3] * q[3]);
        if (!(qlen > 1e-8f)) {
            q[0] = 0.0f; q[1] = 0.0f; q[2] = 0.0f; q[3] = 1.0f;
        }
        const float b0 = decode_colour(r[e->dc0], e->colour_mode);
        const float b1 = decode_colour(r[e->dc1], e->colour_mode);
        const float b2 = decode_colour(r[e->dc2], e->colour_mode);
        shape[j * 4 + 0] = or_encode_quaternion(q);
        shape[j * 4 + 1] = or_pack_halves(half_safe(l0), half_safe(l1));
        shape[j * 4 + 2] = or_pack_halves(half_safe(l2), half_safe(b0));
        shape[j * 4 + 3] = or_pack_halves(half_safe(b1), half_safe(b2));
        if (keep == 0) continue;
        uint32_t word = 0, filled = 0, at = 0;
        for (uint32_t basis = 0; basis < keep; ++basis) {
            for (uint32_t colour = 0; colour < 3; ++colour) {
                const uint32_t field = e->rest_colour_outer ? e->rest_base + colour * e->file_per_colour + basis
                                                            : e->rest_base + basis * 3 + colour;
                float value = decode_byte_unit(r[field], e->rest_mode);
                if (e->flip_yz) value *= kRestFlip[basis < 14u ? basis : 14u];
                word |= (uint32_t)or_float_to_half(half_safe(value)) << (filled * 16);
                if (++filled == 2) {
                    sh[(size_t)j * sh_words + at++] = word;
                    word = 0;
                    filled = 0;
                }
            }
        }
        if (filled == 1) sh[(size_t)j * sh_words + at] = word;
    }
    return kept;
}

/* ====================================================================== */
/* camera and frame (modules/aofx/sdk/include/aofx/Transform.h,            */
/* modules/render/src/Camera.cpp, modules/render/src/FrameParams.h)        */
/* ====================================================================== */

#define AT(m, r, c) (m)[(r) * 4 + (c)]

static void identity(double m[16]) {
    memset(m, 0, 16 * sizeof(double));
    m[0] = m[5] = m[10] = m[15] = 1.0;
}

/* Transform.h:97-110 */
void or_mat_mul(const double a[16], const double b[16], double out[16]) {
    double t[16];
    for (int r = 0; r < 4; ++r) {
        for (int c = 0; c < 4; ++c) {
            double sum = 0.0;
            for (int k = 0; k < 4; ++k) sum += AT(a, r, k) * AT(b, k, c);
            AT(t, r, c) = sum;
        }
    }
    memcpy(out, t, sizeof t);
}

void or_translation(const double t[3], double out[16]) {
    identity(out);
    AT(out, 0, 3) = t[0];
    AT(out, 1, 3) = t[1];
    AT(out, 2, 3) = t[2];
}

void or_scaling(const double s[3], double out[16]) {
    identity(out);
    AT(out, 0, 0) = s[0];
    AT(out, 1, 1) = s[1];
    AT(out, 2, 2) = s[2];
}

static const double kToRadians = 3.14159265358979323846 / 180.0;

void or_rotation_y(double degrees, double out[16]) {
    const double c = cos(degrees * kToRadians), s = sin(degrees * kToRadians);
    identity(out);
    AT(out, 0, 0) = c;
    AT(out, 0, 2) = s;
    AT(out, 2, 0) = -s;
    AT(out, 2, 2) = c;
}

void or_rotation_z(double degrees, double out[16]) {
    const double c = cos(degrees * kToRadians), s = sin(degrees * kToRadians);
    identity(out);
    AT(out, 0, 0) = c;
    AT(out, 0, 1) = -s;
    AT(out, 1, 0) = s;
    AT(out, 1, 1) = c;
}

/* Transform.h:334-359 */
void or_inverse_affine(const double m[16], double out[16]) {
    const double a00 = AT(m, 0, 0), a01 = AT(m, 0, 1), a02 = AT(m, 0, 2);
    const double a10 = AT(m, 1, 0), a11 = AT(m, 1, 1), a12 = AT(m, 1, 2);
    const double a20 = AT(m, 2, 0), a21 = AT(m, 2, 1), a22 = AT(m, 2, 2);
    const double c00 = a11 * a22 - a12 * a21;
    const double c01 = a12 * a20 - a10 * a22;
    const double c02 = a10 * a21 - a11 * a20;
    const double det = a00 * c00 + a01 * c01 + a02 * c02;
    const double inv = fabs(det) > 1e-18 ? 1.0 / det : 0.0;
    double o[16];
    identity(o);
    AT(o, 0, 0) = c00 * inv;
    AT(o, 0, 1) = (a02 * a21 - a01 * a22) * inv;
    AT(o, 0, 2) = (a01 * a12 - a02 * a11) * inv;
    AT(o, 1, 0) = c01 * inv;
    AT(o, 1, 1) = (a00 * a22 - a02 * a20) * inv;
    AT(o, 1, 2) = (a02 * a10 - a00 * a12) * inv;
    AT(o, 2, 0) = c02 * inv;
    AT(o, 2, 1) = (a01 * a20 - a00 * a21) * inv;
    AT(o, 2, 2) = (a00 * a11 - a01 * a10) * inv;
    const double tx = AT(m, 0, 3), ty = AT(m, 1, 3), tz = AT(m, 2, 3);
    AT(o, 0, 3) = (AT(o, 0, 0) * tx + AT(o, 0, 1) * ty + AT(o, 0, 2) * tz) * -1.0;
    AT(o, 1, 3) = (AT(o, 1, 0) * tx + AT(o, 1, 1) * ty + AT(o, 1, 2) * tz) * -1.0;
    AT(o, 2, 3) = (AT(o, 2, 0) * tx + AT(o, 2, 1) * ty + AT(o, 2, 2) * tz) * -1.0;
    memcpy(out, o, sizeof o);
}

static void v_norm(const double a[3], double out[3]) {
    const double l = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    if (l > 1e-12) {
        const double s = 1.0 / l;
        out[0] = a[0] * s; out[1] = a[1] * s; out[2] = a[2] * s;
    } else {
        out[0] = out[1] = out[2] = 0.0;
    }
}

static void v_cross(const double a[3], const double b[3], double out[3]) {
    const double x = a[1] * b[2] - a[2] * b[1];
    const double y = a[2] * b[0] - a[0] * b[2];
    const double z = a[0] * b[1] - a[1] * b[0];
    out[0] = x; out[1] = y; out[2] = z;
}

static double v_len(const double a[3]) { return sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }

/* Transform.h:376-400 with the default -Z look axis */
void or_look_at(const double eye[3], const double target[3], const double up[3], double out[16]) {
    const double d[3] = {target[0] - eye[0], target[1] - eye[1], target[2] - eye[2]};
    double f[3];
    v_norm(d, f);
    if (v_len(f) < 1e-12) {
        or_translation(eye, out);
        return;
    }
    const double z[3] = {f[0] * -1.0, f[1] * -1.0, f[2] * -1.0};
    double x[3], t[3];
    v_cross(up, z, t);
    v_norm(t, x);
    if (v_len(x) < 1e-12) {
        const double zz[3] = {0.0, 0.0, 1.0};
        v_cross(zz, z, t);
        v_norm(t, x);
        if (v_len(x) < 1e-12) {
            x[0] = 1.0; x[1] = 0.0; x[2] = 0.0;
        }
    }
    double y[3];
    v_cross(z, x, y);
    identity(out);
    AT(out, 0, 0) = x[0]; AT(out, 0, 1) = y[0]; AT(out, 0, 2) = z[0]; AT(out, 0, 3) = eye[0];
    AT(out, 1, 0) = x[1]; AT(out, 1, 1) = y[1]; AT(out, 1, 2) = z[1]; AT(out, 1, 3) = eye[1];
    AT(out, 2, 0) = x[2]; AT(out, 2, 1) = y[2]; AT(out, 2, 2) = z[2]; AT(out, 2, 3) = eye[2];
}

/* modules/render/src/Camera.cpp:15-48 */
void or_projection_for(const or_camera* c, uint32_t width, uint32_t height, or_projection* out) {
    double inv[16], flip[16];
    const double s[3] = {1.0, 1.0, -1.0};
    or_inverse_affine(c->camera_to_world, inv);
    or_scaling(s, flip);
    or_mat_mul(flip, inv, out->world_to_view);
    if (fabs(c->window_roll) > 1e-12) {
        double roll[16];
        or_rotation_z(c->window_roll, roll);
        or_mat_mul(roll, out->world_to_view, out->world_to_view);
    }
    const double w = (double)width, h = (double)height;
    const double focal = w * fmax(c->focal, 1e-9) / fmax(c->haperture, 1e-9);
    out->focal_x = focal * c->window_scale[0];
    out->focal_y = focal * c->window_scale[1];
    out->centre_x = w * 0.5 + c->window_translate[0] * w * 0.5;
    out->centre_y = h * 0.5 + c->window_translate[1] * w * 0.5;
    out->near_z = c->near_z;
    out->far_z = c->far_z;
    out->orthographic = c->orthographic;
    out->eye_world[0] = AT(c->camera_to_world, 0, 3);
    out->eye_world[1] = AT(c->camera_to_world, 1, 3);
    out->eye_world[2] = AT(c->camera_to_world, 2, 3);
}

/* FrameParams.h:18-24 */
uint32_t or_bits_for(uint32_t values) {
    uint32_t bits = 1;
    while (bits < 32 && ((uint64_t)1 << bits) < values) ++bits;
    return bits;
}

/* FrameParams.h:32-69 and TileRasterizer.cpp:297-314 */
void or_frame_for(const or_projection* p, const or_settings* s, const double object_to_world[16],
                  const or_edit* edit, uint32_t count, uint32_t base, uint32_t rest_per_colour,
                  uint32_t sh_words, or_frame* f) {
    memset(f, 0, sizeof *f);
    f->width = s->width;
    f->height = s->height;
    f->tile_shift = s->tile_size == 8 ? 3u : s->tile_size == 32 ? 5u : 4u;
    f->tiles_x = (s->width + (1u << f->tile_shift) - 1u) >> f->tile_shift;
    f->tiles_y = (s->height + (1u << f->tile_shift) - 1u) >> f->tile_shift;
    f->focal_x = (float)p->focal_x;
    f->focal_y = (float)p->focal_y;
    f->centre_x = (float)p->centre_x;
    f->centre_y = (float)p->centre_y;
    f->near_z = (float)p->near_z;
    f->far_z = (float)p->far_z;
    f->lim_lo_x = (-0.15f * (float)f->width - f->centre_x) / f->focal_x;
    f->lim_hi_x = (1.15f * (float)f->width - f->centre_x) / f->focal_x;
    f->lim_lo_y = (-0.15f * (float)f->height - f->centre_y) / f->focal_y;
    f->lim_hi_y = (1.15f * (float)f->height - f->centre_y) / f->focal_y;
    f->orthographic = p->orthographic ? 1u : 0u;
    f->antialias = s->antialias ? 1u : 0u;
    double ov[16], inv[16];
    or_mat_mul(p->world_to_view, object_to_world, ov);
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 4; ++c) f->m[r * 4 + c] = (float)AT(ov, r, c);
    or_inverse_affine(object_to_world, inv);
    const double* e = p->eye_world;
    for (int r = 0; r < 3; ++r)
        f->eye[r] = (float)(AT(inv, r, 0) * e[0] + AT(inv, r, 1) * e[1] + AT(inv, r, 2) * e[2] + AT(inv, r, 3));
    f->count = count;
    f->base = base;
    f->rest_per_colour = rest_per_colour;
    f->sh_words = sh_words;
    f->sh_limit = s->max_sh_degree;
    f->tile_bits = or_bits_for(f->tiles_x * f->tiles_y);
    f->depth_mode = s->depth_mode ? 1u : 0u;
    f->depth_threshold = s->depth_threshold;
    memcpy(f->bg, s->background, sizeof f->bg);
    f->linearise = s->linearise ? 1u : 0u;
    f->dilation = s->dilation;
    f->use_filter3d = s->filter3d ? 1u : 0u;
    if (edit) f->edit = *edit;
}

/* ====================================================================== */
/* shared definitions of frame.slang, sh.slang, edit.slang, color.slang    */
/* ====================================================================== */

/* frame.slang:105-107 */
static inline float cutoff_power(float alpha) {
    return fmin_(2.0f * or_logf(fmax_(255.0f * alpha, 1.0001f)), 9.0f);
}

/* frame.slang:113-130 */
static float min_power_over_box(float a, float b, float c, float lox, float loy, float hix, float hiy) {
    if (lox <= 0.0f && hix >= 0.0f && loy <= 0.0f && hiy >= 0.0f) return 0.0f;
    float best = 3.4e38f;
    for (int e = 0; e < 2; ++e) {
        const float x = e == 0 ? lox : hix;
        const float y = c > 1e-12f ? clampf_(-b * x / c, loy, hiy) : loy;
        best = fmin_(best, a * x * x + 2.0f * b * x * y + c * y * y);
    }
    for (int e = 0; e < 2; ++e) {
        const float y = e == 0 ? loy : hiy;
        const float x = a > 1e-12f ? clampf_(-b * y / a, lox, hix) : lox;
        best = fmin_(best, a * x * x + 2.0f * b * x * y + c * y * y);
    }
    return best;
}

/* frame.slang:133-138 */
static inline int tile_touched(const or_frame* f, float cx, float cy, float a, float b, float c,
                               float cutoff, uint32_t tx, uint32_t ty) {
    const uint32_t side = 1u << f->tile_shift;
    const float fx = (float)(tx * side) + 0.5f;
    const float fy = (float)(ty * side) + 0.5f;
    const uint32_t ex = (tx + 1u) * side < f->width ? (tx + 1u) * side : f->width;
    const uint32_t ey = (ty + 1u) * side < f->height ? (ty + 1u) * side : f->height;
    const float lx = (float)(ex - 1u) + 0.5f;
    const float ly = (float)(ey - 1u) + 0.5f;
    return min_power_over_box(a, b, c, fx - cx, fy - cy, lx - cx, ly - cy) <= cutoff;
}

/* frame.slang:142-146 */
static inline void footprint_extent(float a, float b, float c, float cutoff, float* ex, float* ey) {
    const float det = a * c - b * b;
    const float inv = det > 1e-20f ? 1.0f / det : 0.0f;
    *ex = sqrtf(fmax_(cutoff * c * inv, 0.0f));
    *ey = sqrtf(fmax_(cutoff * a * inv, 0.0f));
}

/* frame.slang:149-161 */
static inline void tile_rect(const or_frame* f, float cx, float cy, float ex, float ey, uint32_t r[4]) {
    const float x1 = cx - ex, y1 = cy - ey, x2 = cx + ex, y2 = cy + ey;
    if (x2 < 0.0f || y2 < 0.0f || x1 >= (float)f->width || y1 >= (float)f->height) {
        r[0] = 1; r[1] = 1; r[2] = 0; r[3] = 0;
        return;
    }
    r[0] = (uint32_t)fmax_(x1, 0.0f) >> f->tile_shift;
    r[1] = (uint32_t)fmax_(y1, 0.0f) >> f->tile_shift;
    r[2] = (uint32_t)fmin_(x2, (float)f->width - 1.0f) >> f->tile_shift;
    r[3] = (uint32_t)fmin_(y2, (float)f->height - 1.0f) >> f->tile_shift;
}

/* sh.slang:8-13 */
static const float kC1 = 0.4886025119029199f;
static const float kC2[5] = {1.0925484305920792f, -1.0925484305920792f, 0.31539156525252005f,
                             -1.0925484305920792f, 0.5462742152960396f};
static const float kC3[7] = {-0.5900435899266435f, 2.890611442640554f, -0.4570457994644658f,
                             0.3731763325901154f, -0.4570457994644658f, 1.445305721320277f,
                             -0.5900435899266435f};

static inline float half_at(const uint32_t* row, uint32_t index) {
    const uint32_t w = row[index / 2];
    return (index & 1u) == 0u ? low_half(w) : high_half(w);
}

/* sh.slang:25-50: the view-dependent part, channel ch. */
static float evaluate_rest(const uint32_t* row, uint32_t degree, float x, float y, float z, int ch) {
    if (degree == 0) return 0.0f;
#define B(k) half_at(row, (k) * 3u + (uint32_t)ch)
    float result = -kC1 * y * B(0) + kC1 * z * B(1) - kC1 * x * B(2);
    if (degree < 2) return result;
    const float xx = x * x, yy = y * y, zz = z * z;
    const float xy = x * y, yz = y * z, xz = x * z;
    result = result + (kC2[0] * xy * B(3) + kC2[1] * yz * B(4) + kC2[2] * (2.0f * zz - xx - yy) * B(5) +
                       kC2[3] * xz * B(6) + kC2[4] * (xx - yy) * B(7));
    if (degree < 3) return result;
    result = result + (kC3[0] * y * (3.0f * xx - yy) * B(8) + kC3[1] * xy * z * B(9) +
                       kC3[2] * y * (4.0f * zz - xx - yy) * B(10) +
                       kC3[3] * z * (2.0f * zz - 3.0f * xx - 3.0f * yy) * B(11) +
                       kC3[4] * x * (4.0f * zz - xx - yy) * B(12) + kC3[5] * z * (xx - yy) * B(13) +
                       kC3[6] * x * (xx - 3.0f * yy) * B(14));
#undef B
    return result;
}

/* sh.slang:89-91 */
static inline uint32_t degree_of(uint32_t rest) {
    return rest >= 15 ? 3 : rest >= 8 ? 2 : rest >= 3 ? 1 : 0;
}

/* color.slang:7-10 */
float or_srgb_to_linear(float c) {
    c = fmax_(c, 0.0f);
    return c <= 0.04045f ? c / 12.92f : or_powf((c + 0.055f) / 1.055f, 2.4f);
}

typedef struct { int removed, graded; float opacity; } edit_fate;

/* edit.slang:46-78 */
static edit_fate edit_fate_of(const or_edit* e, float px, float py, float pz, float opacity,
                              float sx, float sy, float sz) {
    edit_fate fate = {0, 0, opacity};
    if (e->active == 0) return fate;
    if (e->min_opacity > 0.0f && opacity < e->min_opacity) {
        fate.removed = 1;
        return fate;
    }
    if (e->max_scale > 0.0f && fmax_(fmax_(fabsf(sx), fabsf(sy)), fabsf(sz)) > e->max_scale) {
        fate.removed = 1;
        return fate;
    }
    const float rx = px - e->centre[0], ry = py - e->centre[1], rz = pz - e->centre[2];
    int inside = e->shape == 1 ? ((rx * rx + ry * ry) + rz * rz) <= e->size[0] * e->size[0]
                               : (fabsf(rx) <= e->size[0] && fabsf(ry) <= e->size[1] && fabsf(rz) <= e->size[2]);
    if (e->invert != 0) inside = !inside;
    if ((e->mode == 0 && !inside) || (e->mode == 1 && inside)) {
        fate.removed = 1;
        return fate;
    }
    if (inside) {
        fate.graded = 1;
        fate.opacity = opacity * e->opacity;
        fate.removed = !(fate.opacity >= 1.0f / 255.0f);
    }
    return fate;
}

/* edit.slang:81-85 */
static void edit_grade(const or_edit* e, float c[3]) {
    const float r = c[0] * e->tint[0] * e->brightness;
    const float g = c[1] * e->tint[1] * e->brightness;
    const float b = c[2] * e->tint[2] * e->brightness;
    const float luma = (r * 0.2126f + g * 0.7152f) + b * 0.0722f;
    /* lerp(a, b, t) = a + (b - a) * t */
    c[0] = fmax_(luma + (r - luma) * e->saturation, 0.0f);
    c[1] = fmax_(luma + (g - luma) * e->saturation, 0.0f);
    c[2] = fmax_(luma + (b - luma) * e->saturation, 0.0f);
}

/* Mip-Splatting's 3D filter (not in the reference): scale^2 + f^2 per axis and
 * opacity * sqrt(prod s^2 / prod (s^2 + f^2)). */
static inline void filter3d_apply(float f, float* sx, float* sy, float* sz, float* opacity) {
    const float f2 = f * f;
    const float s2x = *sx * *sx, s2y = *sy * *sy, s2z = *sz * *sz;
    const float a2x = s2x + f2, a2y = s2y + f2, a2z = s2z + f2;
    const float det1 = (s2x * s2y) * s2z;
    const float det2 = (a2x * a2y) * a2z;
    *opacity = *opacity * sqrtf(det1 / det2);
    *sx = sqrtf(a2x);
    *sy = sqrtf(a2y);
    *sz = sqrtf(a2z);
}

/* ====================================================================== */
/* splatProject                                                            */
/* ====================================================================== */

static void cull(uint32_t slot, uint32_t* rects, uint32_t* touched, uint32_t* visible, uint32_t* keys) {
    if (touched) touched[slot] = 0;
    if (visible) visible[slot] = 0;
    if (keys) keys[slot] = 0x00FFFFFFu;
    if (rects) {
        rects[slot * 4 + 0] = 1; rects[slot * 4 + 1] = 1;
        rects[slot * 4 + 2] = 0; rects[slot * 4 + 3] = 0;
    }
}

static void project_one(const or_frame* f, uint32_t i, const float* positions, const uint32_t* shape,
                        const uint32_t* shbuf, const float* filter3d, or_proj_record* proj,
                        uint32_t* rects, uint32_t* touched_out, uint32_t* visible, uint32_t* keys) {
    const uint32_t slot = f->base + i;
    const float* p = positions + (size_t)i * 4;
    const uint32_t* row = shape + (size_t)i * 4;
    float sx = or_expf(low_half(row[1]));
    float sy = or_expf(high_half(row[1]));
    float sz = or_expf(low_half(row[2]));
    const edit_fate fate = edit_fate_of(&f->edit, p[0], p[1], p[2], p[3], sx, sy, sz);
    if (fate.removed) {
        cull(slot, rects, touched_out, visible, keys);
        return;
    }
    float opacity = fate.opacity;
    if (f->use_filter3d) filter3d_apply(filter3d[i], &sx, &sy, &sz, &opacity);

    const float* m = f->m;
    const float tx_ = m[0] * p[0] + m[1] * p[1] + m[2] * p[2] + m[3];
    const float ty_ = m[4] * p[0] + m[5] * p[1] + m[6] * p[2] + m[7];
    const float tz_ = m[8] * p[0] + m[9] * p[1] + m[10] * p[2] + m[11];
    if (!(tz_ > f->near_z) || tz_ > f->far_z) {
        cull(slot, rects, touched_out, visible, keys);
        return;
    }
    float q[4];
    or_decode_quaternion(row[0], q);
    const float x = q[0], y = q[1], z = q[2], w = q[3];
    const float r[3][3] = {
        {1.0f - 2.0f * (y * y + z * z), 2.0f * (x * y - w * z), 2.0f * (x * z + w * y)},
        {2.0f * (x * y + w * z), 1.0f - 2.0f * (x * x + z * z), 2.0f * (y * z - w * x)},
        {2.0f * (x * z - w * y), 2.0f * (y * z + w * x), 1.0f - 2.0f * (x * x + y * y)}};
    const float sc[3] = {sx, sy, sz};
    float mm[3][3];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) mm[a][b] = r[a][b] * sc[b];

    float jx[3], jy[3], cx, cy;
    if (f->orthographic) {
        jx[0] = f->focal_x; jx[1] = 0.0f; jx[2] = 0.0f;
        jy[0] = 0.0f; jy[1] = f->focal_y; jy[2] = 0.0f;
        cx = f->focal_x * tx_ + f->centre_x;
        cy = f->focal_y * ty_ + f->centre_y;
    } else {
        const float invz = 1.0f / tz_;
        cx = f->focal_x * tx_ * invz + f->centre_x;
        cy = f->focal_y * ty_ * invz + f->centre_y;
        const float llx = (-0.15f * (float)f->width - f->centre_x) / f->focal_x;
        const float lhx = (1.15f * (float)f->width - f->centre_x) / f->focal_x;
        const float lly = (-0.15f * (float)f->height - f->centre_y) / f->focal_y;
        const float lhy = (1.15f * (float)f->height - f->centre_y) / f->focal_y;
        const float ctx = clampf_(tx_ * invz, llx, lhx) * tz_;
        const float cty = clampf_(ty_ * invz, lly, lhy) * tz_;
        jx[0] = f->focal_x * invz; jx[1] = 0.0f; jx[2] = -f->focal_x * ctx * invz * invz;
        jy[0] = 0.0f; jy[1] = f->focal_y * invz; jy[2] = -f->focal_y * cty * invz * invz;
    }
    /* row vectors times W, then times M */
    float t3x[3], t3y[3], ax[3], ay[3];
    for (int c = 0; c < 3; ++c) {
        t3x[c] = jx[0] * m[0 * 4 + c] + jx[1] * m[1 * 4 + c] + jx[2] * m[2 * 4 + c];
        t3y[c] = jy[0] * m[0 * 4 + c] + jy[1] * m[1 * 4 + c] + jy[2] * m[2 * 4 + c];
    }
    for (int c = 0; c < 3; ++c) {
        ax[c] = t3x[0] * mm[0][c] + t3x[1] * mm[1][c] + t3x[2] * mm[2][c];
        ay[c] = t3y[0] * mm[0][c] + t3y[1] * mm[1][c] + t3y[2] * mm[2][c];
    }
    float a = ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2];
    const float b = ax[0] * ay[0] + ax[1] * ay[1] + ax[2] * ay[2];
    float c = ay[0] * ay[0] + ay[1] * ay[1] + ay[2] * ay[2];

    const float det_before = a * c - b * b;
    a = a + f->dilation;
    c = c + f->dilation;
    const float det = a * c - b * b;
    if (!(det > 1e-12f)) {
        cull(slot, rects, touched_out, visible, keys);
        return;
    }
    float alpha = opacity;
    if (f->antialias) alpha = alpha * sqrtf(fmax_(det_before, 0.0f) / det);
    if (!(alpha >= 1.0f / 255.0f)) {
        cull(slot, rects, touched_out, visible, keys);
        return;
    }
    const float ca = c / det, cb = -b / det, cc = a / det;
    const float cutoff = cutoff_power(alpha);
    float ex, ey;
    footprint_extent(ca, cb, cc, cutoff, &ex, &ey);
    uint32_t rect[4];
    tile_rect(f, cx, cy, ex, ey, rect);
    if (rect[0] > rect[2] || rect[1] > rect[3]) {
        cull(slot, rects, touched_out, visible, keys);
        return;
    }
    uint32_t touched = 0;
    for (uint32_t tyy = rect[1]; tyy <= rect[3]; ++tyy)
        for (uint32_t txx = rect[0]; txx <= rect[2]; ++txx)
            touched += (uint32_t)tile_touched(f, cx, cy, ca, cb, cc, cutoff, txx, tyy);
    if (touched == 0) {
        cull(slot, rects, touched_out, visible, keys);
        return;
    }
    const uint32_t degree = degree_of(f->rest_per_colour) < f->sh_limit ? degree_of(f->rest_per_colour) : f->sh_limit;
    const float dxo = p[0] - f->eye[0], dyo = p[1] - f->eye[1], dzo = p[2] - f->eye[2];
    const float len = sqrtf((dxo * dxo + dyo * dyo) + dzo * dzo);
    const float dx = dxo / len, dy = dyo / len, dz = dzo / len;
    const float base[3] = {high_half(row[2]), low_half(row[3]), high_half(row[3])};
    const uint32_t* shrow = shbuf + (size_t)i * f->sh_words;
    float colour[3];
    for (int ch = 0; ch < 3; ++ch)
        colour[ch] = fmax_(base[ch] + evaluate_rest(shrow, degree, dx, dy, dz, ch), 0.0f);
    if (fate.graded) edit_grade(&f->edit, colour);

    or_proj_record* out = proj + slot;
    out->screen[0] = cx; out->screen[1] = cy; out->screen[2] = tz_; out->screen[3] = cutoff;
    out->conic[0] = ca; out->conic[1] = cb; out->conic[2] = cc; out->conic[3] = alpha;
    out->colour[0] = colour[0]; out->colour[1] = colour[1]; out->colour[2] = colour[2]; out->colour[3] = 0.0f;
    if (rects) memcpy(rects + (size_t)slot * 4, rect, sizeof rect);
    if (touched_out) touched_out[slot] = touched;
    if (visible) visible[slot] = 1;
    if (keys) keys[slot] = f2u(tz_) >> 7;
}

void or_project(const or_frame* f, const float* positions, const uint32_t* shape, const uint32_t* sh,
                const float* filter3d, or_proj_record* proj, uint32_t* rects, uint32_t* touched,
                uint32_t* visible, uint32_t* depth_keys, int threads) {
    (void)threads;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 4096) OR_THREADS(threads)
#endif
    for (int64_t i = 0; i < (int64_t)f->count; ++i)
        project_one(f, (uint32_t)i, positions, shape, sh, filter3d, proj, rects, touched, visible, depth_keys);
}

/* pointRadius (shaders/lrt/points/point_frame.slang:39-44): the disc's radius
 * in pixels, never below half a pixel. */
static float point_radius(const or_frame* f, const or_point_style* st, float view_z) {
    const float r = st->size_mode == 1 ? st->size * 0.5f
                                       : (f->orthographic ? st->size * 0.5f * f->focal_x
                                                          : st->size * 0.5f * f->focal_x / view_z);
    return fmax_(r, 0.5f);
}

/* pointsProject (shaders/lrt/splat/points_project.slang:33-86): a disc record
 * with the conic of its radius, cutoff 1, alpha 1 and colour.w = 1; the tile
 * rectangle of the radius itself (:52-53) and the same exact tile count. */
static void points_project_one(const or_frame* f, const or_point_style* st, uint32_t i, const float* positions,
                               const uint32_t* colours, or_proj_record* proj, uint32_t* rects,
                               uint32_t* touched_out, uint32_t* visible, uint32_t* keys) {
    const uint32_t slot = f->base + i;
    const float* p = positions + (size_t)i * 4;
    const float* m = f->m;
    const float tx_ = m[0] * p[0] + m[1] * p[1] + m[2] * p[2] + m[3];
    const float ty_ = m[4] * p[0] + m[5] * p[1] + m[6] * p[2] + m[7];
    const float tz_ = m[8] * p[0] + m[9] * p[1] + m[10] * p[2] + m[11];
    if (!(tz_ > f->near_z) || tz_ > f->far_z) {
        cull(slot, rects, touched_out, visible, keys);
        return;
    }
    const float radius = point_radius(f, st, tz_);
    const float invz = f->orthographic ? 1.0f : 1.0f / tz_;
    const float cx = f->focal_x * tx_ * invz + f->centre_x;
    const float cy = f->focal_y * ty_ * invz + f->centre_y;
    const float k = 1.0f / (radius * radius);
    const float cutoff = 1.0f;
    uint32_t rect[4];
    tile_rect(f, cx, cy, radius, radius, rect);
    if (rect[0] > rect[2] || rect[1] > rect[3]) {
        cull(slot, rects, touched_out, visible, keys);
        return;
    }
    uint32_t touched = 0;
    for (uint32_t tyy = rect[1]; tyy <= rect[3]; ++tyy)
        for (uint32_t txx = rect[0]; txx <= rect[2]; ++txx)
            touched += (uint32_t)tile_touched(f, cx, cy, k, 0.0f, k, cutoff, txx, tyy);
    if (touched == 0) {
        cull(slot, rects, touched_out, visible, keys);
        return;
    }
    float rgb[3] = {st->colour[0], st->colour[1], st->colour[2]};
    if (!st->constant_colour) {
        rgb[0] = low_half(colours[(size_t)i * 2]);
        rgb[1] = high_half(colours[(size_t)i * 2]);
        rgb[2] = low_half(colours[(size_t)i * 2 + 1]);
    }
    or_proj_record* out = proj + slot;
    out->screen[0] = cx; out->screen[1] = cy; out->screen[2] = tz_; out->screen[3] = cutoff;
    out->conic[0] = k; out->conic[1] = 0.0f; out->conic[2] = k; out->conic[3] = 1.0f;
    out->colour[0] = rgb[0]; out->colour[1] = rgb[1]; out->colour[2] = rgb[2]; out->colour[3] = 1.0f;
    if (rects) memcpy(rects + (size_t)slot * 4, rect, sizeof rect);
    if (touched_out) touched_out[slot] = touched;
    if (visible) visible[slot] = 1;
    if (keys) keys[slot] = f2u(tz_) >> 7;
}

void or_points_project(const or_frame* f, const or_point_style* style, const float* positions,
                       const uint32_t* colours, or_proj_record* proj, uint32_t* rects, uint32_t* touched,
                       uint32_t* visible, uint32_t* depth_keys, int threads) {
    (void)threads;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 4096) OR_THREADS(threads)
#endif
    for (int64_t i = 0; i < (int64_t)f->count; ++i)
        points_project_one(f, style, (uint32_t)i, positions, colours, proj, rects, touched, visible, depth_keys);
}

/* ====================================================================== */
/* scan, sort, emit, ranges                                                */
/* ====================================================================== */

/* PrefixSum::apply semantics (modules/gpu/src/algo/PrefixSum.cpp:58-88). */
uint64_t or_scan_exclusive(const uint32_t* in, uint32_t* out, uint32_t n) {
    uint64_t run = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t v = in[i];
        out[i] = (uint32_t)run;
        run += v;
    }
    return run;
}

/* RadixSort::sort semantics: stable LSD, 8 bits a pass (RadixSort.cpp:80-158). */
void or_radix_sort(uint64_t* keys, uint32_t* values, uint32_t n, uint32_t key_bits) {
    if (n < 2) return;
    uint64_t* k2 = (uint64_t*)malloc((size_t)n * sizeof *k2);
    uint32_t* v2 = (uint32_t*)malloc((size_t)n * sizeof *v2);
    uint64_t* ks = keys;
    uint32_t* vs = values;
    uint64_t* kd = k2;
    uint32_t* vd = v2;
    for (uint32_t shift = 0; shift < key_bits; shift += 8) {
        const uint32_t width = key_bits - shift < 8 ? key_bits - shift : 8;
        const uint64_t mask = ((uint64_t)1 << width) - 1;
        uint64_t count[257] = {0};
        for (uint32_t i = 0; i < n; ++i) count[((ks[i] >> shift) & mask) + 1]++;
        for (int d = 0; d < 256; ++d) count[d + 1] += count[d];
        for (uint32_t i = 0; i < n; ++i) {
            const uint64_t at = count[(ks[i] >> shift) & mask]++;
            kd[at] = ks[i];
            vd[at] = vs[i];
        }
        uint64_t* tk = ks; ks = kd; kd = tk;
        uint32_t* tv = vs; vs = vd; vd = tv;
    }
    if (ks != keys) {
        memcpy(keys, ks, (size_t)n * sizeof *keys);
        memcpy(values, vs, (size_t)n * sizeof *values);
    }
    free(k2);
    free(v2);
}

/* ====================================================================== */
/* splatBlend                                                              */
/* ====================================================================== */

/* blendPixel (shaders/lrt/splat/splat_blend.slang:23-108). With composite set
 * (splatBlendComposite, :116-120): the walk ends at the first record at or
 * behind the under-layer's depth (:52-54), and an opaque disc that reaches the
 * pixel ends it with its colour behind the splats (:63-68); without a disc the
 * under-layer is what lies behind (:83-87). under_colour/under_depth may be
 * NULL (no under-layer, params.hasUnder == 0). */
static void blend_pixel(const or_frame* f, uint32_t px, uint32_t py, const uint32_t* ranges,
                        const uint32_t* pair_splats, const or_proj_record* proj, const float* under_colour,
                        const float* under_depth, int composite, float* colour, float* depth,
                        uint32_t* last_count, float* final_t) {
    const uint32_t tile = (py >> f->tile_shift) * f->tiles_x + (px >> f->tile_shift);
    const uint32_t start = ranges[tile * 2], end = ranges[tile * 2 + 1];
    const float pxf = (float)px + 0.5f, pyf = (float)py + 0.5f;
    const size_t at = (size_t)py * f->width + px;
    const int under = composite && under_depth != NULL;
    const float under_z = under ? under_depth[at] : 0.0f;
    const int under_hit = under && under_z > 0.0f;
    float r = 0.0f, g = 0.0f, b = 0.0f, t = 1.0f, depth_sum = 0.0f, depth_at = -1.0f;
    int opaque_hit = 0;
    float opaque[4] = {0.0f, 0.0f, 0.0f, 0.0f}, opaque_z = 0.0f;
    uint32_t last = 0;
    for (uint32_t k = start; k < end; ++k) {
        const or_proj_record* rec = proj + pair_splats[k];
        if (under_hit && rec->screen[2] >= under_z) break;
        const float dx = pxf - rec->screen[0], dy = pyf - rec->screen[1];
        const float power = rec->conic[0] * dx * dx + 2.0f * rec->conic[1] * dx * dy + rec->conic[2] * dy * dy;
        if (power > rec->screen[3]) continue;
        if (composite && rec->colour[3] > 0.5f) {
            opaque_hit = 1;
            opaque[0] = rec->colour[0]; opaque[1] = rec->colour[1]; opaque[2] = rec->colour[2]; opaque[3] = 1.0f;
            opaque_z = rec->screen[2];
            break;
        }
        const float alpha = fmin_(0.99f, rec->conic[3] * or_expf(-0.5f * power));
        if (alpha < 1.0f / 255.0f) continue;
        const float weight = t * alpha;
        r = r + weight * rec->colour[0];
        g = g + weight * rec->colour[1];
        b = b + weight * rec->colour[2];
        depth_sum = depth_sum + weight * rec->screen[2];
        t = t * (1.0f - alpha);
        last = k - start + 1;
        if (depth_at < 0.0f && 1.0f - t >= f->depth_threshold) depth_at = rec->screen[2];
        if (t < 1e-4f) break;
    }
    if (!opaque_hit && under_hit) {
        opaque_hit = 1;
        memcpy(opaque, under_colour + at * 4, sizeof opaque);
        opaque_z = under_z;
    }
    const float coverage = 1.0f - t;
    if (f->linearise && coverage > 1e-6f) {
        r = or_srgb_to_linear(r / coverage) * coverage;
        g = or_srgb_to_linear(g / coverage) * coverage;
        b = or_srgb_to_linear(b / coverage) * coverage;
    }
    const float* behind = f->bg;
    if (opaque_hit) {
        behind = opaque;
        depth_sum = depth_sum + t * opaque_z;
        if (depth_at < 0.0f) depth_at = opaque_z;
    }
    const float total = opaque_hit ? 1.0f : coverage;
    colour[at * 4 + 0] = r + t * behind[0];
    colour[at * 4 + 1] = g + t * behind[1];
    colour[at * 4 + 2] = b + t * behind[2];
    colour[at * 4 + 3] = coverage + t * behind[3];
    if (depth) {
        depth[at] = f->depth_mode == 0 ? (total > 1e-6f ? depth_sum / total : 0.0f) : fmax_(depth_at, 0.0f);
    }
    if (last_count) last_count[at] = last;
    if (final_t) final_t[at] = t;
}

void or_blend(const or_frame* f, const uint32_t* ranges, const uint32_t* pair_splats,
              const or_proj_record* proj, float* colour, float* depth, uint32_t* last_count,
              float* final_t, int threads) {
    (void)threads;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) OR_THREADS(threads)
#endif
    for (int64_t py = 0; py < (int64_t)f->height; ++py)
        for (uint32_t px = 0; px < f->width; ++px)
            blend_pixel(f, px, (uint32_t)py, ranges, pair_splats, proj, NULL, NULL, 0, colour, depth, last_count,
                        final_t);
}

void or_blend_composite(const or_frame* f, const uint32_t* ranges, const uint32_t* pair_splats,
                        const or_proj_record* proj, const float* under_colour, const float* under_depth,
                        float* colour, float* depth, int threads) {
    (void)threads;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) OR_THREADS(threads)
#endif
    for (int64_t py = 0; py < (int64_t)f->height; ++py)
        for (uint32_t px = 0; px < f->width; ++px)
            blend_pixel(f, px, (uint32_t)py, ranges, pair_splats, proj, under_colour, under_depth, 1, colour, depth,
                        NULL, NULL);
}

/* ====================================================================== */
/* TileRasterizer::render (TileRasterizer.cpp:188-484), splat-only          */
/* ====================================================================== */

/* Splat instances' slots first, then the point instances' (TileRasterizer.cpp:
 * 296-379); the composite blend when there are points or an under-layer
 * (:464). */
static uint32_t raster_frame(const or_frame* frames, const or_cloud* clouds, uint32_t instances,
                             const or_frame* point_frames, const or_points* points, uint32_t point_instances,
                             const float* under_colour, const float* under_depth, or_proj_record* proj,
                             uint32_t* touched, uint32_t* depth_keys, uint32_t* pair_tiles, uint32_t* pair_splats,
                             uint32_t pair_capacity, uint32_t* ranges, float* colour, float* depth,
                             uint32_t* last_count, float* final_t, or_stats* stats, int threads) {
    uint32_t total = 0;
    for (uint32_t k = 0; k < instances; ++k) total += frames[k].count;
    for (uint32_t k = 0; k < point_instances; ++k) total += point_frames[k].count;
    const int composite = point_instances > 0 || under_depth != NULL;
    const uint32_t all = total > 0 ? total : 1;
    const or_frame* f0 = instances > 0 || point_instances == 0 ? &frames[0] : &point_frames[0];   /* shared fields */
    const uint32_t tiles = f0->tiles_x * f0->tiles_y;
    uint32_t* rects = (uint32_t*)calloc((size_t)all * 4, sizeof(uint32_t));
    uint32_t* visible = (uint32_t*)calloc(all, sizeof(uint32_t));
    uint32_t* vis_off = (uint32_t*)calloc(all, sizeof(uint32_t));
    uint32_t* touch_off = (uint32_t*)calloc(all, sizeof(uint32_t));
    int own_touched = touched == NULL, own_keys = depth_keys == NULL, own_proj = proj == NULL;
    if (own_touched) touched = (uint32_t*)calloc(all, sizeof(uint32_t));
    if (own_keys) depth_keys = (uint32_t*)calloc(all, sizeof(uint32_t));
    if (own_proj) proj = (or_proj_record*)calloc(all, sizeof(or_proj_record));

    for (uint32_t k = 0; k < instances; ++k)
        or_project(&frames[k], clouds[k].positions, clouds[k].shape, clouds[k].sh, clouds[k].filter3d, proj,
                   rects, touched, visible, depth_keys, threads);
    for (uint32_t k = 0; k < point_instances; ++k)
        or_points_project(&point_frames[k], &points[k].style, points[k].positions, points[k].colours, proj, rects,
                          touched, visible, depth_keys, threads);

    const uint32_t visible_count = (uint32_t)or_scan_exclusive(visible, vis_off, total);
    const uint32_t pairs = (uint32_t)or_scan_exclusive(touched, touch_off, total);

    /* splat_compact.slang:16-26, then the 24-bit stable depth sort (:411) */
    uint64_t* sort_keys = (uint64_t*)calloc(visible_count > 0 ? visible_count : 1, sizeof(uint64_t));
    uint32_t* order = (uint32_t*)calloc(visible_count > 0 ? visible_count : 1, sizeof(uint32_t));
    for (uint32_t i = 0; i < total; ++i) {
        if (!visible[i]) continue;
        sort_keys[vis_off[i]] = depth_keys[i];
        order[vis_off[i]] = i;
    }
    if (visible_count > 1) or_radix_sort(sort_keys, order, visible_count, 24);

    /* splat_gather_counts + scan + splat_emit (:418-439) */
    uint32_t* offsets = (uint32_t*)calloc(visible_count > 0 ? visible_count : 1, sizeof(uint32_t));
    uint32_t* sorted_counts = (uint32_t*)calloc(visible_count > 0 ? visible_count : 1, sizeof(uint32_t));
    for (uint32_t k = 0; k < visible_count; ++k) sorted_counts[k] = touched[order[k]];
    or_scan_exclusive(sorted_counts, offsets, visible_count);
    uint64_t* ptiles = (uint64_t*)calloc(pairs > 0 ? pairs : 1, sizeof(uint64_t));
    uint32_t* psplats = (uint32_t*)calloc(pairs > 0 ? pairs : 1, sizeof(uint32_t));
    for (uint32_t k = 0; k < visible_count; ++k) {
        const uint32_t s = order[k];
        const uint32_t expected = touched[s];
        if (expected == 0) continue;
        const uint32_t first = offsets[k];
        const or_proj_record* rec = proj + s;
        const uint32_t* rc = rects + (size_t)s * 4;
        uint32_t written = 0;
        for (uint32_t ty = rc[1]; ty <= rc[3] && written < expected; ++ty) {
            for (uint32_t tx = rc[0]; tx <= rc[2] && written < expected; ++tx) {
                if (!tile_touched(f0, rec->screen[0], rec->screen[1], rec->conic[0], rec->conic[1], rec->conic[2],
                                  rec->screen[3], tx, ty))
                    continue;
                const uint32_t at = first + written;
                if (at < pairs) {
                    ptiles[at] = ty * f0->tiles_x + tx;
                    psplats[at] = s;
                }
                written += 1;
            }
        }
    }
    /* the stable tile sort on bitsFor(T) bits (:445) */
    if (pairs > 1) or_radix_sort(ptiles, psplats, pairs, or_bits_for(tiles));

    /* splat_tiles_clear + splat_ranges (:447-459) */
    uint32_t* rng = ranges ? ranges : (uint32_t*)calloc((size_t)tiles * 2, sizeof(uint32_t));
    memset(rng, 0, (size_t)tiles * 2 * sizeof(uint32_t));
    for (uint32_t i = 0; i < pairs; ++i) {
        const uint32_t t = (uint32_t)ptiles[i];
        if (i == 0 || (uint32_t)ptiles[i - 1] != t) rng[t * 2] = i;
        if (i + 1 == pairs || (uint32_t)ptiles[i + 1] != t) rng[t * 2 + 1] = i + 1;
    }
    if (colour && composite) {
        or_blend_composite(f0, rng, psplats, proj, under_colour, under_depth, colour, depth, threads);
    } else if (colour) {
        or_blend(f0, rng, psplats, proj, colour, depth, last_count, final_t, threads);
    }

    const uint32_t keep = pairs < pair_capacity ? pairs : pair_capacity;
    for (uint32_t i = 0; i < keep; ++i) {
        if (pair_tiles) pair_tiles[i] = (uint32_t)ptiles[i];
        if (pair_splats) pair_splats[i] = psplats[i];
    }
    if (stats) {
        stats->splats = total;
        stats->visible = visible_count;
        stats->pairs = pairs;
    }
    if (!ranges) free(rng);
    free(ptiles); free(psplats); free(offsets); free(sorted_counts); free(sort_keys); free(order);
    free(rects); free(visible); free(vis_off); free(touch_off);
    if (own_touched) free(touched);
    if (own_keys) free(depth_keys);
    if (own_proj) free(proj);
    return pairs;
}

uint32_t or_raster(const or_frame* frames, const or_cloud* clouds, uint32_t instances,
                   or_proj_record* proj, uint32_t* touched, uint32_t* depth_keys,
                   uint32_t* pair_tiles, uint32_t* pair_splats, uint32_t pair_capacity,
                   uint32_t* ranges, float* colour, float* depth, uint32_t* last_count,
                   float* final_t, or_stats* stats, int threads) {
    return raster_frame(frames, clouds, instances, NULL, NULL, 0, NULL, NULL, proj, touched, depth_keys, pair_tiles,
                        pair_splats, pair_capacity, ranges, colour, depth, last_count, final_t, stats, threads);
}

uint32_t or_raster_composite(const or_frame* frames, const or_cloud* clouds, uint32_t instances,
                             const or_frame* point_frames, const or_points* points, uint32_t point_instances,
                             const float* under_colour, const float* under_depth, or_proj_record* proj,
                             uint32_t* touched, uint32_t* depth_keys, uint32_t* pair_tiles, uint32_t* pair_splats,
                             uint32_t pair_capacity, uint32_t* ranges, float* colour, float* depth,
                             or_stats* stats, int threads) {
    return raster_frame(frames, clouds, instances, point_frames, points, point_instances, under_colour, under_depth,
                        proj, touched, depth_keys, pair_tiles, pair_splats, pair_capacity, ranges, colour, depth,
                        NULL, NULL, stats, threads);
}

/* ====================================================================== */
/* The brute-force ground truth                                            */
/* ====================================================================== */

typedef struct { float screen[4], conic[4], colour[3]; int disc; } ref_f32;
typedef struct { double screen[4], conic[4], colour[3]; } ref_f64;

/* reference_project.slang:31-105 in FP32. screen[2] < 0 marks a culled splat. */
static void ref_project_f32(const or_frame* f, const or_cloud* cl, uint32_t i, ref_f32* out) {
    out->screen[2] = -1.0f;
    const float* p = cl->positions + (size_t)i * 4;
    const uint32_t* row = cl->shape + (size_t)i * 4;
    float sx = or_expf(low_half(row[1])), sy = or_expf(high_half(row[1])), sz = or_expf(low_half(row[2]));
    const edit_fate fate = edit_fate_of(&f->edit, p[0], p[1], p[2], p[3], sx, sy, sz);
    if (fate.removed) return;
    float opacity = fate.opacity;
    if (f->use_filter3d) filter3d_apply(cl->filter3d[i], &sx, &sy, &sz, &opacity);
    const float* m = f->m;
    const float v[3] = {m[0] * p[0] + m[1] * p[1] + m[2] * p[2] + m[3],
                        m[4] * p[0] + m[5] * p[1] + m[6] * p[2] + m[7],
                        m[8] * p[0] + m[9] * p[1] + m[10] * p[2] + m[11]};
    if (!(v[2] > f->near_z) || v[2] > f->far_z) return;
    float q[4];
    or_decode_quaternion(row[0], q);
    const float x = q[0], y = q[1], z = q[2], w = q[3];
    const float r[3][3] = {
        {1.0f - 2.0f * (y * y + z * z), 2.0f * (x * y - w * z), 2.0f * (x * z + w * y)},
        {2.0f * (x * y + w * z), 1.0f - 2.0f * (x * x + z * z), 2.0f * (y * z - w * x)},
        {2.0f * (x * z - w * y), 2.0f * (y * z + w * x), 1.0f - 2.0f * (x * x + y * y)}};
    const float s[3] = {sx, sy, sz};
    float rs[3][3], cov3[3][3], tmp[3][3], cv[3][3];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) rs[a][b] = (r[a][0] * (b == 0 ? s[0] : 0.0f) + r[a][1] * (b == 1 ? s[1] : 0.0f)) +
                                              r[a][2] * (b == 2 ? s[2] : 0.0f);
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) cov3[a][b] = (rs[a][0] * rs[b][0] + rs[a][1] * rs[b][1]) + rs[a][2] * rs[b][2];
    /* covView = W (cov3 W^T) */
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
            tmp[a][b] = (cov3[a][0] * m[b * 4 + 0] + cov3[a][1] * m[b * 4 + 1]) + cov3[a][2] * m[b * 4 + 2];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) cv[a][b] = (m[a * 4 + 0] * tmp[0][b] + m[a * 4 + 1] * tmp[1][b]) + m[a * 4 + 2] * tmp[2][b];
    float j[3][3] = {{0}};
    float cx, cy;
    if (f->orthographic) {
        cx = f->focal_x * v[0] + f->centre_x;
        cy = f->focal_y * v[1] + f->centre_y;
        j[0][0] = f->focal_x;
        j[1][1] = f->focal_y;
    } else {
        cx = f->focal_x * v[0] / v[2] + f->centre_x;
        cy = f->focal_y * v[1] / v[2] + f->centre_y;
        const float lox = (-0.15f * (float)f->width - f->centre_x) / f->focal_x;
        const float hix = (1.15f * (float)f->width - f->centre_x) / f->focal_x;
        const float loy = (-0.15f * (float)f->height - f->centre_y) / f->focal_y;
        const float hiy = (1.15f * (float)f->height - f->centre_y) / f->focal_y;
        const float ccx = clampf_(v[0] / v[2], lox, hix) * v[2];
        const float ccy = clampf_(v[1] / v[2], loy, hiy) * v[2];
        j[0][0] = f->focal_x / v[2];
        j[0][2] = -f->focal_x * ccx / (v[2] * v[2]);
        j[1][1] = f->focal_y / v[2];
        j[1][2] = -f->focal_y * ccy / (v[2] * v[2]);
    }
    /* cov2 = J (covView J^T) */
    float t2[3][3], c2[3][3];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) t2[a][b] = (cv[a][0] * j[b][0] + cv[a][1] * j[b][1]) + cv[a][2] * j[b][2];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) c2[a][b] = (j[a][0] * t2[0][b] + j[a][1] * t2[1][b]) + j[a][2] * t2[2][b];
    const float before = c2[0][0] * c2[1][1] - c2[0][1] * c2[0][1];
    const float a = c2[0][0] + f->dilation, b = c2[0][1], c = c2[1][1] + f->dilation;
    const float det = a * c - b * b;
    if (!(det > 1e-12f)) return;
    float alpha = opacity;
    if (f->antialias) alpha = alpha * sqrtf(fmax_(before, 0.0f) / det);
    if (!(alpha >= 1.0f / 255.0f)) return;
    const uint32_t degree = degree_of(f->rest_per_colour) < f->sh_limit ? degree_of(f->rest_per_colour) : f->sh_limit;
    const float dxo = p[0] - f->eye[0], dyo = p[1] - f->eye[1], dzo = p[2] - f->eye[2];
    const float len = sqrtf((dxo * dxo + dyo * dyo) + dzo * dzo);
    const float base[3] = {high_half(row[2]), low_half(row[3]), high_half(row[3])};
    const uint32_t* shrow = cl->sh + (size_t)i * f->sh_words;
    float colour[3];
    for (int ch = 0; ch < 3; ++ch)
        colour[ch] = fmax_(base[ch] + evaluate_rest(shrow, degree, dxo / len, dyo / len, dzo / len, ch), 0.0f);
    if (fate.graded) edit_grade(&f->edit, colour);
    out->screen[0] = cx; out->screen[1] = cy; out->screen[2] = v[2]; out->screen[3] = cutoff_power(alpha);
    out->conic[0] = c / det; out->conic[1] = -b / det; out->conic[2] = a / det; out->conic[3] = alpha;
    memcpy(out->colour, colour, sizeof colour);
}

static double srgb_to_linear_d(double c) {
    c = c > 0.0 ? c : 0.0;
    return c <= 0.04045 ? c / 12.92 : pow((c + 0.055) / 1.055, 2.4);
}

static double half_d(uint32_t w, int high) { return (double)(high ? high_half(w) : low_half(w)); }

/* reference_project.slang:31-105 in FP64 (exp, log, sqrt from libm in double),
 * from decoded inputs: centre p, opacity, scales s (after any 3D filter), the
 * quaternion q (normalised here), the base colour and the rest coefficients
 * (rgb per basis). Returns 0 for a culled splat; the colour before any edit
 * grade is left in out->colour. */
static int ref_project_core_f64(const or_frame* f, const double p[3], double opacity, const double s[3],
                                const double qin[4], const double base[3], const double* coeff, uint32_t degree,
                                ref_f64* out) {
    out->screen[2] = -1.0;
    const float* mf = f->m;
    double m[12];
    for (int k = 0; k < 12; ++k) m[k] = mf[k];
    const double v[3] = {m[0] * p[0] + m[1] * p[1] + m[2] * p[2] + m[3],
                         m[4] * p[0] + m[5] * p[1] + m[6] * p[2] + m[7],
                         m[8] * p[0] + m[9] * p[1] + m[10] * p[2] + m[11]};
    if (!(v[2] > (double)f->near_z) || v[2] > (double)f->far_z) return 0;
    double q[4] = {qin[0], qin[1], qin[2], qin[3]};
    const double ql = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    if (ql > 0.0) {
        for (int k = 0; k < 4; ++k) q[k] /= ql;
    } else {
        q[0] = q[1] = q[2] = 0.0;
        q[3] = 1.0;
    }
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
        for (int b = 0; b < 3; ++b) tmp[a][b] = cov3[a][0] * m[b * 4 + 0] + cov3[a][1] * m[b * 4 + 1] + cov3[a][2] * m[b * 4 + 2];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) cv[a][b] = m[a * 4 + 0] * tmp[0][b] + m[a * 4 + 1] * tmp[1][b] + m[a * 4 + 2] * tmp[2][b];
    double j[3][3] = {{0}};
    double cx, cy;
    const double fx = f->focal_x, fy = f->focal_y, ccx0 = f->centre_x, ccy0 = f->centre_y;
    if (f->orthographic) {
        cx = fx * v[0] + ccx0;
        cy = fy * v[1] + ccy0;
        j[0][0] = fx;
        j[1][1] = fy;
    } else {
        cx = fx * v[0] / v[2] + ccx0;
        cy = fy * v[1] / v[2] + ccy0;
        const double lox = (-0.15 * (double)f->width - ccx0) / fx;
        const double hix = (1.15 * (double)f->width - ccx0) / fx;
        const double loy = (-0.15 * (double)f->height - ccy0) / fy;
        const double hiy = (1.15 * (double)f->height - ccy0) / fy;
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
    const double dil = f->dilation;
    const double before = c2[0][0] * c2[1][1] - c2[0][1] * c2[0][1];
    const double a = c2[0][0] + dil, b = c2[0][1], c = c2[1][1] + dil;
    const double det = a * c - b * b;
    if (!(det > 1e-12)) return 0;
    double alpha = opacity;
    if (f->antialias) alpha *= sqrt((before > 0.0 ? before : 0.0) / det);
    if (!(alpha >= 1.0 / 255.0)) return 0;
    const double cut = 2.0 * log(255.0 * alpha > 1.0001 ? 255.0 * alpha : 1.0001);
    const double dxo = p[0] - f->eye[0], dyo = p[1] - f->eye[1], dzo = p[2] - f->eye[2];
    const double len = sqrt(dxo * dxo + dyo * dyo + dzo * dzo);
    const double dx = dxo / len, dy = dyo / len, dz = dzo / len;
    for (int ch = 0; ch < 3; ++ch) {
        double rest = 0.0;
        if (degree >= 1) {
#define BD(k) coeff[(k) * 3 + ch]
            rest = -0.4886025119029199 * dy * BD(0) + 0.4886025119029199 * dz * BD(1) - 0.4886025119029199 * dx * BD(2);
            if (degree >= 2) {
                const double xx = dx * dx, yy = dy * dy, zz = dz * dz;
                rest += 1.0925484305920792 * dx * dy * BD(3) - 1.0925484305920792 * dy * dz * BD(4) +
                        0.31539156525252005 * (2.0 * zz - xx - yy) * BD(5) - 1.0925484305920792 * dx * dz * BD(6) +
                        0.5462742152960396 * (xx - yy) * BD(7);
                if (degree >= 3) {
                    rest += -0.5900435899266435 * dy * (3.0 * xx - yy) * BD(8) + 2.890611442640554 * dx * dy * dz * BD(9) -
                            0.4570457994644658 * dy * (4.0 * zz - xx - yy) * BD(10) +
                            0.3731763325901154 * dz * (2.0 * zz - 3.0 * xx - 3.0 * yy) * BD(11) -
                            0.4570457994644658 * dx * (4.0 * zz - xx - yy) * BD(12) +
                            1.445305721320277 * dz * (xx - yy) * BD(13) -
                            0.5900435899266435 * dx * (xx - 3.0 * yy) * BD(14);
                }
            }
#undef BD
        }
        const double v0 = base[ch] + rest;
        out->colour[ch] = v0 > 0.0 ? v0 : 0.0;
    }
    out->screen[0] = cx; out->screen[1] = cy; out->screen[2] = v[2]; out->screen[3] = cut < 9.0 ? cut : 9.0;
    out->conic[0] = c / det; out->conic[1] = -b / det; out->conic[2] = a / det; out->conic[3] = alpha;
    return 1;
}

static void filter3d_f64(double fl, double s[3], double* opacity) {
    const double f2 = fl * fl;
    double det1 = 1.0, det2 = 1.0;
    for (int k = 0; k < 3; ++k) {
        det1 *= s[k] * s[k];
        det2 *= s[k] * s[k] + f2;
        s[k] = sqrt(s[k] * s[k] + f2);
    }
    *opacity *= sqrt(det1 / det2);
}

/* The packed cloud's splat i in FP64: its decoded fp16/10-bit inputs, then the core. */
static void ref_project_f64(const or_frame* f, const or_cloud* cl, uint32_t i, ref_f64* out) {
    out->screen[2] = -1.0;
    const float* pf = cl->positions + (size_t)i * 4;
    const double p[3] = {pf[0], pf[1], pf[2]};
    const uint32_t* row = cl->shape + (size_t)i * 4;
    double s[3] = {exp(half_d(row[1], 0)), exp(half_d(row[1], 1)), exp(half_d(row[2], 0))};
    const edit_fate fate = edit_fate_of(&f->edit, pf[0], pf[1], pf[2], pf[3], (float)s[0], (float)s[1], (float)s[2]);
    if (fate.removed) return;
    double opacity = fate.graded ? (double)pf[3] * (double)f->edit.opacity : (double)pf[3];
    if (f->use_filter3d) filter3d_f64(cl->filter3d[i], s, &opacity);
    float qf[4];
    or_decode_quaternion(row[0], qf);
    /* The stored quaternion's decode is part of the input; the core normalises
     * it in double. */
    const double q[4] = {qf[0], qf[1], qf[2], qf[3]};
    const uint32_t degree = degree_of(f->rest_per_colour) < f->sh_limit ? degree_of(f->rest_per_colour) : f->sh_limit;
    const double base[3] = {half_d(row[2], 1), half_d(row[3], 0), half_d(row[3], 1)};
    const uint32_t* shrow = cl->sh + (size_t)i * f->sh_words;
    double coeff[45] = {0};
    const uint32_t used = degree == 0 ? 0u : degree == 1 ? 9u : degree == 2 ? 24u : 45u;
    for (uint32_t k = 0; k < used; ++k) coeff[k] = (double)half_at(shrow, k);
    if (!ref_project_core_f64(f, p, opacity, s, q, base, coeff, degree, out)) return;
    if (fate.graded) {
        double* colour = out->colour;
        const double tr = colour[0] * f->edit.tint[0] * f->edit.brightness;
        const double tg = colour[1] * f->edit.tint[1] * f->edit.brightness;
        const double tb = colour[2] * f->edit.tint[2] * f->edit.brightness;
        const double luma = tr * 0.2126 + tg * 0.7152 + tb * 0.0722;
        const double t3[3] = {tr, tg, tb};
        for (int k = 0; k < 3; ++k) {
            const double g = luma + (t3[k] - luma) * f->edit.saturation;
            colour[k] = g > 0.0 ? g : 0.0;
        }
    }
}

/* The FP32-parameter cloud's splat i in FP64: scales exp(log s), opacity
 * sigmoid(logit), base colour 0.5 + C0 dc, then the core. */
static void ref_project_diff_f64(const or_frame* f, const or_diff_cloud* c, uint32_t i, ref_f64* out) {
    out->screen[2] = -1.0;
    const double p[3] = {c->means[i * 3 + 0], c->means[i * 3 + 1], c->means[i * 3 + 2]};
    double s[3] = {exp((double)c->log_scales[i * 3 + 0]), exp((double)c->log_scales[i * 3 + 1]),
                   exp((double)c->log_scales[i * 3 + 2])};
    double opacity = 1.0 / (1.0 + exp(-(double)c->logits[i]));
    if (f->use_filter3d) filter3d_f64(c->filter3d[i], s, &opacity);
    const double q[4] = {c->quats[i * 4 + 0], c->quats[i * 4 + 1], c->quats[i * 4 + 2], c->quats[i * 4 + 3]};
    const uint32_t degree = degree_of(c->rest_per_colour) < f->sh_limit ? degree_of(c->rest_per_colour) : f->sh_limit;
    const size_t stride = (size_t)(1 + c->rest_per_colour) * 3;
    const float* row = c->sh + (size_t)i * stride;
    const double base[3] = {0.5 + 0.28209479177387814 * row[0], 0.5 + 0.28209479177387814 * row[1],
                            0.5 + 0.28209479177387814 * row[2]};
    double coeff[45] = {0};
    const uint32_t used = degree == 0 ? 0u : degree == 1 ? 9u : degree == 2 ? 24u : 45u;
    for (uint32_t k = 0; k < used; ++k) coeff[k] = row[3 + k];
    (void)ref_project_core_f64(f, p, opacity, s, q, base, coeff, degree, out);
}

typedef struct { double z; uint32_t index; } depth_entry;

/* The rasteriser's own tie rule at its own resolution: the 24-bit key
 * bits(float z) >> 7 (splat_project.slang:216-220), then the index. */
static double keyed_depth(double z) {
    return (double)(f2u((float)z) >> 7);
}

static int by_depth(const void* a, const void* b) {
    const depth_entry* x = (const depth_entry*)a;
    const depth_entry* y = (const depth_entry*)b;
    if (x->z < y->z) return -1;
    if (x->z > y->z) return 1;
    return x->index < y->index ? -1 : x->index > y->index ? 1 : 0;
}

/* referencePointsProject (shaders/lrt/reference/reference_points_project.slang:
 * 20-42): the disc without tiles, its centre by division. */
static void ref_points_project_f32(const or_frame* f, const or_points* pts, uint32_t i, ref_f32* out) {
    out->screen[2] = -1.0f;
    out->disc = 1;
    const float* p = pts->positions + (size_t)i * 4;
    const float* m = f->m;
    const float v[3] = {m[0] * p[0] + m[1] * p[1] + m[2] * p[2] + m[3],
                        m[4] * p[0] + m[5] * p[1] + m[6] * p[2] + m[7],
                        m[8] * p[0] + m[9] * p[1] + m[10] * p[2] + m[11]};
    if (!(v[2] > f->near_z) || v[2] > f->far_z) return;
    const or_point_style* st = &pts->style;
    const float r = st->size_mode == 1 ? fmax_(st->size * 0.5f, 0.5f)
                                       : fmax_(f->orthographic ? st->size * 0.5f * f->focal_x
                                                               : st->size * 0.5f * f->focal_x / v[2], 0.5f);
    out->screen[0] = f->orthographic ? f->focal_x * v[0] + f->centre_x : f->focal_x * v[0] / v[2] + f->centre_x;
    out->screen[1] = f->orthographic ? f->focal_y * v[1] + f->centre_y : f->focal_y * v[1] / v[2] + f->centre_y;
    out->screen[2] = v[2];
    out->screen[3] = 1.0f;
    out->conic[0] = 1.0f / (r * r); out->conic[1] = 0.0f; out->conic[2] = 1.0f / (r * r); out->conic[3] = 1.0f;
    if (st->constant_colour) {
        memcpy(out->colour, st->colour, sizeof out->colour);
    } else {
        out->colour[0] = low_half(pts->colours[(size_t)i * 2]);
        out->colour[1] = high_half(pts->colours[(size_t)i * 2]);
        out->colour[2] = low_half(pts->colours[(size_t)i * 2 + 1]);
    }
}

/* The walk of reference_blend.slang:24-119, over every splat in (z, index)
 * order: the same sequence its per-pixel insertion sort builds. */
static void reference_f32(const or_frame* frames, const or_cloud* clouds, uint32_t instances,
                          const or_frame* point_frames, const or_points* points, uint32_t point_instances,
                          const uint32_t* px, const uint32_t* py, uint32_t npix, float* colour, float* depth,
                          int keyed, int threads) {
    (void)threads;
    uint32_t total = 0;
    for (uint32_t k = 0; k < instances; ++k) total += frames[k].count;
    for (uint32_t k = 0; k < point_instances; ++k) total += point_frames[k].count;
    ref_f32* rec = (ref_f32*)calloc(total > 0 ? total : 1, sizeof(ref_f32));
    for (uint32_t k = 0; k < instances; ++k) {
#ifdef _OPENMP
#pragma omp parallel for schedule(static) OR_THREADS(threads)
#endif
        for (int64_t i = 0; i < (int64_t)frames[k].count; ++i)
            ref_project_f32(&frames[k], &clouds[k], (uint32_t)i, rec + frames[k].base + i);
    }
    for (uint32_t k = 0; k < point_instances; ++k) {
        for (int64_t i = 0; i < (int64_t)point_frames[k].count; ++i)
            ref_points_project_f32(&point_frames[k], &points[k], (uint32_t)i, rec + point_frames[k].base + i);
    }
    depth_entry* order = (depth_entry*)malloc((total > 0 ? total : 1) * sizeof(depth_entry));
    uint32_t live = 0;
    for (uint32_t s = 0; s < total; ++s)
        if (rec[s].screen[2] >= 0.0f) {
            order[live].z = keyed ? keyed_depth(rec[s].screen[2]) : rec[s].screen[2];
            order[live].index = s;
            live++;
        }
    qsort(order, live, sizeof *order, by_depth);
    ref_f32* sorted = (ref_f32*)malloc((live > 0 ? live : 1) * sizeof(ref_f32));
    for (uint32_t k = 0; k < live; ++k) sorted[k] = rec[order[k].index];
    const or_frame* f = instances > 0 || point_instances == 0 ? &frames[0] : &point_frames[0];
    const uint32_t count = px ? npix : f->width * f->height;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 64) OR_THREADS(threads)
#endif
    for (int64_t n = 0; n < (int64_t)count; ++n) {
        const uint32_t x = px ? px[n] : (uint32_t)(n % f->width);
        const uint32_t y = px ? py[n] : (uint32_t)(n / f->width);
        const float pxf = (float)x + 0.5f, pyf = (float)y + 0.5f;
        float r = 0.0f, g = 0.0f, b = 0.0f, t = 1.0f, dsum = 0.0f, dat = -1.0f;
        const ref_f32* disc = NULL;
        for (uint32_t k = 0; k < live; ++k) {
            const ref_f32* s = sorted + k;
            const float dx = pxf - s->screen[0], dy = pyf - s->screen[1];
            const float power = s->conic[0] * dx * dx + 2.0f * s->conic[1] * dx * dy + s->conic[2] * dy * dy;
            if (power > s->screen[3]) continue;
            if (s->disc) {   /* alpha 1, and the walk ends there (reference_blend.slang:50-51,83-88) */
                disc = s;
                break;
            }
            const float alpha = fmin_(0.99f, s->conic[3] * or_expf(-0.5f * power));
            if (alpha < 1.0f / 255.0f) continue;
            const float weight = t * alpha;
            r = r + weight * s->colour[0];
            g = g + weight * s->colour[1];
            b = b + weight * s->colour[2];
            dsum = dsum + weight * s->screen[2];
            t = t * (1.0f - alpha);
            if (dat < 0.0f && 1.0f - t >= f->depth_threshold) dat = s->screen[2];
            if (t < 1e-4f) break;
        }
        const float coverage = 1.0f - t;
        if (f->linearise && coverage > 1e-6f) {
            r = or_srgb_to_linear(r / coverage) * coverage;
            g = or_srgb_to_linear(g / coverage) * coverage;
            b = or_srgb_to_linear(b / coverage) * coverage;
        }
        const float behind[4] = {disc ? disc->colour[0] : f->bg[0], disc ? disc->colour[1] : f->bg[1],
                                 disc ? disc->colour[2] : f->bg[2], disc ? 1.0f : f->bg[3]};
        if (disc) {
            dsum = dsum + t * disc->screen[2];
            if (dat < 0.0f) dat = disc->screen[2];
        }
        const float total_w = disc ? 1.0f : coverage;
        colour[n * 4 + 0] = r + t * behind[0];
        colour[n * 4 + 1] = g + t * behind[1];
        colour[n * 4 + 2] = b + t * behind[2];
        colour[n * 4 + 3] = coverage + t * behind[3];
        if (depth) depth[n] = f->depth_mode == 0 ? (total_w > 1e-6f ? dsum / total_w : 0.0f) : fmax_(dat, 0.0f);
    }
    free(sorted);
    free(order);
    free(rec);
}

void or_reference_render_f32(const or_frame* frames, const or_cloud* clouds, uint32_t instances,
                             const uint32_t* px, const uint32_t* py, uint32_t npix, float* colour,
                             float* depth, int keyed, int threads) {
    reference_f32(frames, clouds, instances, NULL, NULL, 0, px, py, npix, colour, depth, keyed, threads);
}

void or_reference_render_points_f32(const or_frame* frames, const or_cloud* clouds, uint32_t instances,
                                    const or_frame* point_frames, const or_points* points,
                                    uint32_t point_instances, float* colour, float* depth, int threads) {
    reference_f32(frames, clouds, instances, point_frames, points, point_instances, NULL, NULL, 0, colour, depth, 0,
                  threads);
}

/* The FP64 walk over every projected splat in (z, index) order (or the keyed
 * order), as or_reference_render_f32's. */
static void blend_f64(const or_frame* frames, ref_f64* rec, uint32_t total, const uint32_t* px, const uint32_t* py,
                      uint32_t npix, double* colour, double* depth, int keyed, int threads) {
    (void)threads;
    depth_entry* order = (depth_entry*)malloc((total > 0 ? total : 1) * sizeof(depth_entry));
    uint32_t live = 0;
    for (uint32_t s = 0; s < total; ++s)
        if (rec[s].screen[2] >= 0.0) {
            order[live].z = keyed ? keyed_depth(rec[s].screen[2]) : rec[s].screen[2];
            order[live].index = s;
            live++;
        }
    qsort(order, live, sizeof *order, by_depth);
    ref_f64* sorted = (ref_f64*)malloc((live > 0 ? live : 1) * sizeof(ref_f64));
    for (uint32_t k = 0; k < live; ++k) sorted[k] = rec[order[k].index];
    const or_frame* f = &frames[0];
    const uint32_t count = px ? npix : f->width * f->height;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 64) OR_THREADS(threads)
#endif
    for (int64_t n = 0; n < (int64_t)count; ++n) {
        const uint32_t x = px ? px[n] : (uint32_t)(n % f->width);
        const uint32_t y = px ? py[n] : (uint32_t)(n / f->width);
        const double pxf = (double)x + 0.5, pyf = (double)y + 0.5;
        double r = 0.0, g = 0.0, b = 0.0, t = 1.0, dsum = 0.0, dat = -1.0;
        for (uint32_t k = 0; k < live; ++k) {
            const ref_f64* s = sorted + k;
            const double dx = pxf - s->screen[0], dy = pyf - s->screen[1];
            const double power = s->conic[0] * dx * dx + 2.0 * s->conic[1] * dx * dy + s->conic[2] * dy * dy;
            if (power > s->screen[3]) continue;
            double alpha = s->conic[3] * exp(-0.5 * power);
            alpha = alpha < 0.99 ? alpha : 0.99;
            if (alpha < 1.0 / 255.0) continue;
            const double weight = t * alpha;
            r += weight * s->colour[0];
            g += weight * s->colour[1];
            b += weight * s->colour[2];
            dsum += weight * s->screen[2];
            t *= 1.0 - alpha;
            if (dat < 0.0 && 1.0 - t >= (double)f->depth_threshold) dat = s->screen[2];
            if (t < 1e-4) break;
        }
        const double coverage = 1.0 - t;
        if (f->linearise && coverage > 1e-6) {
            r = srgb_to_linear_d(r / coverage) * coverage;
            g = srgb_to_linear_d(g / coverage) * coverage;
            b = srgb_to_linear_d(b / coverage) * coverage;
        }
        colour[n * 4 + 0] = r + t * f->bg[0];
        colour[n * 4 + 1] = g + t * f->bg[1];
        colour[n * 4 + 2] = b + t * f->bg[2];
        colour[n * 4 + 3] = coverage + t * f->bg[3];
        if (depth) depth[n] = f->depth_mode == 0 ? (coverage > 1e-6 ? dsum / coverage : 0.0) : (dat > 0.0 ? dat : 0.0);
    }
    free(sorted);
    free(order);
}


void or_reference_render_f64(const or_frame* frames, const or_cloud* clouds, uint32_t instances,
                             const uint32_t* px, const uint32_t* py, uint32_t npix, double* colour,
                             double* depth, int keyed, int threads) {
    (void)threads;
    uint32_t total = 0;
    for (uint32_t k = 0; k < instances; ++k) total += frames[k].count;
    ref_f64* rec = (ref_f64*)calloc(total > 0 ? total : 1, sizeof(ref_f64));
    for (uint32_t k = 0; k < instances; ++k) {
#ifdef _OPENMP
#pragma omp parallel for schedule(static) OR_THREADS(threads)
#endif
        for (int64_t i = 0; i < (int64_t)frames[k].count; ++i)
            ref_project_f64(&frames[k], &clouds[k], (uint32_t)i, rec + frames[k].base + i);
    }
    blend_f64(frames, rec, total, px, py, npix, colour, depth, keyed, threads);
    free(rec);
}

void or_diff_render_f64(const or_frame* frame, const or_diff_cloud* cloud, const uint32_t* px, const uint32_t* py,
                        uint32_t npix, double* colour, double* depth, int keyed, int threads) {
    (void)threads;
    const uint32_t total = cloud->count;
    ref_f64* rec = (ref_f64*)calloc(total > 0 ? total : 1, sizeof(ref_f64));
#ifdef _OPENMP
#pragma omp parallel for schedule(static) OR_THREADS(threads)
#endif
    for (int64_t i = 0; i < (int64_t)total; ++i) ref_project_diff_f64(frame, cloud, (uint32_t)i, rec + i);
    blend_f64(frame, rec, total, px, py, npix, colour, depth, keyed, threads);
    free(rec);
}

/* ====================================================================== */
/* compareImages                                                           */
/* ====================================================================== */

static uint32_t code_of(float linear) {
    const float c = clampf_(linear, 0.0f, 1.0f);
    const float e = c <= 0.0031308f ? c * 12.92f : 1.055f * powf(c, 1.0f / 2.4f) - 0.055f;
    return (uint32_t)lrintf(e * 255.0f);
}

void or_compare_images(const float* a, const float* b, uint32_t width, uint32_t height,
                       uint32_t* p99, uint32_t* max, uint64_t* over2) {
    uint64_t hist[256] = {0};
    const uint64_t pixels = (uint64_t)width * height;
    for (uint64_t i = 0; i < pixels; ++i) {
        uint32_t worst = 0;
        for (int c = 0; c < 3; ++c) {
            const int d = (int)code_of(a[i * 4 + c]) - (int)code_of(b[i * 4 + c]);
            const uint32_t ad = (uint32_t)(d < 0 ? -d : d);
            if (ad > worst) worst = ad;
        }
        const int da = (int)lrintf(clampf_(a[i * 4 + 3], 0.0f, 1.0f) * 255.0f) -
                       (int)lrintf(clampf_(b[i * 4 + 3], 0.0f, 1.0f) * 255.0f);
        const uint32_t ada = (uint32_t)(da < 0 ? -da : da);
        if (ada > worst) worst = ada;
        hist[worst]++;
    }
    uint64_t total = 0, o2 = 0;
    uint32_t mx = 0;
    for (uint32_t k = 0; k < 256; ++k) {
        total += hist[k];
        if (hist[k] > 0) mx = k;
        if (k > 2) o2 += hist[k];
    }
    const uint64_t at = total - total / 100;
    uint64_t run = 0;
    uint32_t p = 0;
    for (uint32_t k = 0; k < 256; ++k) {
        run += hist[k];
        if (run >= at) {
            p = k;
            break;
        }
    }
    *p99 = p;
    *max = mx;
    *over2 = o2;
}

/* ====================================================================== */
/* Levels of detail (modules/lod, shaders/lrt/lod)                         */
/* ====================================================================== */

/* morton.slang:7-34 */
static uint32_t spread_bits(uint32_t v) {
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

static uint32_t compact_bits(uint32_t v) {
    v &= 0x09249249u;
    v = (v | (v >> 2)) & 0x030C30C3u;
    v = (v | (v >> 4)) & 0x0300F00Fu;
    v = (v | (v >> 8)) & 0x030000FFu;
    v = (v | (v >> 16)) & 0x000003FFu;
    return v;
}

static uint32_t quantise_unit(float u) {
    const float s = u > 0.0f ? (u < 1.0f ? u : 1.0f) : 0.0f;   /* saturate */
    const uint32_t q = (uint32_t)(s * 1024.0f);
    return q < 1023u ? q : 1023u;
}

uint32_t or_morton30(float ux, float uy, float uz) {
    return (spread_bits(quantise_unit(ux)) << 2) | (spread_bits(quantise_unit(uy)) << 1) |
           spread_bits(quantise_unit(uz));
}

/* lodMorton (lod_morton.slang:12-21) */
void or_lod_morton(const float* positions, uint32_t n, const float lo[3], float extent, uint64_t* keys,
                   uint32_t* values) {
    const float e = fmax_(extent, 1e-20f);
    for (uint32_t i = 0; i < n; ++i) {
        const float* p = positions + (size_t)i * 4;
        keys[i] = or_morton30((p[0] - lo[0]) / e, (p[1] - lo[1]) / e, (p[2] - lo[2]) / e);
        values[i] = i;
    }
}

/* lodBoundaries + PrefixSum + lodGroups (lod_boundaries.slang:12-21,
 * lod_groups.slang:19-30): each splat's group at `level`, each group's first
 * splat and cell code (top 3 level bits of the 30-bit code). Returns the groups. */
uint32_t or_lod_groups(const uint64_t* keys, uint32_t n, uint32_t level, uint32_t* group, uint32_t* starts,
                       uint32_t* cells) {
    const uint32_t shift = 3u * (10u - level);
    uint32_t groups = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t code = (uint32_t)keys[i];
        const int boundary = i == 0 || (code >> shift) != ((uint32_t)keys[i - 1] >> shift);
        if (boundary) {
            if (starts) starts[groups] = i;
            if (cells) cells[groups] = code >> shift;
            groups++;
        }
        if (group) group[i] = groups - 1;
    }
    return groups;
}

/* lodLeafMoments (lod_leaf_moments.slang:21-64): per group, in splat order,
 * W, S1, S2 (xx xy xz yy yz zz), base colour, then keep * 3 rest halves. */
void or_lod_leaf_moments(const float* positions, const uint32_t* shape, const uint32_t* sh, uint32_t n,
                         const uint32_t* starts, uint32_t groups, uint32_t keep, uint32_t sh_words, float* moments) {
    const uint32_t stride = 13u + keep * 3u;
    for (uint32_t g = 0; g < groups; ++g) {
        const uint32_t first = starts[g], end = g + 1 < groups ? starts[g + 1] : n;
        float* m = moments + (size_t)g * stride;
        for (uint32_t k = 0; k < stride; ++k) m[k] = 0.0f;
        for (uint32_t i = first; i < end; ++i) {
            const float* p = positions + (size_t)i * 4;
            const uint32_t* row = shape + (size_t)i * 4;
            const float s[3] = {or_expf(low_half(row[1])), or_expf(high_half(row[1])), or_expf(low_half(row[2]))};
            float q[4];
            or_decode_quaternion(row[0], q);
            const float x = q[0], y = q[1], z = q[2], w = q[3];
            const float r[9] = {1.0f - 2.0f * (y * y + z * z), 2.0f * (x * y - w * z), 2.0f * (x * z + w * y),
                                2.0f * (x * y + w * z), 1.0f - 2.0f * (x * x + z * z), 2.0f * (y * z - w * x),
                                2.0f * (x * z - w * y), 2.0f * (y * z + w * x), 1.0f - 2.0f * (x * x + y * y)};
            float mm[9], cov[9];
            for (int a = 0; a < 3; ++a)
                for (int c = 0; c < 3; ++c) mm[a * 3 + c] = r[a * 3 + c] * s[c];
            for (int a = 0; a < 3; ++a)
                for (int c = 0; c < 3; ++c)
                    cov[a * 3 + c] = (mm[a * 3] * mm[c * 3] + mm[a * 3 + 1] * mm[c * 3 + 1]) + mm[a * 3 + 2] * mm[c * 3 + 2];
            const float smallest = fmin_(fmin_(s[0], s[1]), s[2]);
            const float area = s[0] * s[1] * s[2] / fmax_(smallest, 1e-20f);
            const float wt = p[3] * area;
            m[0] += wt;
            m[1] += wt * p[0];
            m[2] += wt * p[1];
            m[3] += wt * p[2];
            m[4] += wt * (cov[0] + p[0] * p[0]);
            m[5] += wt * (cov[1] + p[0] * p[1]);
            m[6] += wt * (cov[2] + p[0] * p[2]);
            m[7] += wt * (cov[4] + p[1] * p[1]);
            m[8] += wt * (cov[5] + p[1] * p[2]);
            m[9] += wt * (cov[8] + p[2] * p[2]);
            m[10] += wt * high_half(row[2]);
            m[11] += wt * low_half(row[3]);
            m[12] += wt * high_half(row[3]);
            for (uint32_t h = 0; h < keep * 3; ++h) m[13 + h] += wt * half_at(sh + (size_t)i * sh_words, h);
        }
    }
}

/* lodMergeMoments (lod_merge_moments.slang:17-32) */
void or_lod_merge_moments(const uint32_t* starts, uint32_t groups, const uint32_t* fine_group, uint32_t fine_groups,
                          const float* fine_moments, uint32_t stride, float* moments) {
    for (uint32_t g = 0; g < groups; ++g) {
        const uint32_t first = fine_group[starts[g]];
        const uint32_t end = g + 1 < groups ? fine_group[starts[g + 1]] : fine_groups;
        for (uint32_t k = 0; k < stride; ++k) {
            float sum = 0.0f;
            for (uint32_t h = first; h < end; ++h) sum += fine_moments[(size_t)h * stride + k];
            moments[(size_t)g * stride + k] = sum;
        }
    }
}

typedef struct { float a[3][3]; } m3;

static m3 m3_mul(const m3* x, const m3* y) {
    m3 r;
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) r.a[i][k] = (x->a[i][0] * y->a[0][k] + x->a[i][1] * y->a[1][k]) + x->a[i][2] * y->a[2][k];
    return r;
}

static m3 m3_transposed(const m3* x) {
    m3 r;
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) r.a[i][k] = x->a[k][i];
    return r;
}

static float sign_of(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }

/* jacobi (lod_finalize.slang:16-41): six cyclic sweeps over (0,1), (0,2), (1,2). */
static void jacobi3(m3* a, m3* v) {
    const m3 id = {{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}};
    *v = id;
    for (int sweep = 0; sweep < 6; ++sweep) {
        for (int pq = 0; pq < 3; ++pq) {
            const int p = pq == 2 ? 1 : 0;
            const int q = pq == 0 ? 1 : 2;
            const float apq = a->a[p][q];
            if (fabsf(apq) < 1e-30f) continue;
            const float theta = (a->a[q][q] - a->a[p][p]) / (2.0f * apq);
            const float t = sign_of(theta + 1e-30f) / (fabsf(theta) + sqrtf(theta * theta + 1.0f));
            const float c = 1.0f / sqrtf(t * t + 1.0f);
            const float s = t * c;
            m3 j = id;
            j.a[p][p] = c;
            j.a[q][q] = c;
            j.a[p][q] = s;
            j.a[q][p] = -s;
            const m3 aj = m3_mul(a, &j);
            const m3 jt = m3_transposed(&j);
            *a = m3_mul(&jt, &aj);
            *v = m3_mul(v, &j);
        }
    }
}

/* quaternionOfAxes (packing.slang:76-93) */
static void quaternion_of_axes(const m3* m, float q[4]) {
    const float trace = m->a[0][0] + m->a[1][1] + m->a[2][2];
    if (trace > 0.0f) {
        const float s = 0.5f / sqrtf(trace + 1.0f);
        q[0] = (m->a[2][1] - m->a[1][2]) * s;
        q[1] = (m->a[0][2] - m->a[2][0]) * s;
        q[2] = (m->a[1][0] - m->a[0][1]) * s;
        q[3] = 0.25f / s;
    } else if (m->a[0][0] > m->a[1][1] && m->a[0][0] > m->a[2][2]) {
        const float s = 2.0f * sqrtf(1.0f + m->a[0][0] - m->a[1][1] - m->a[2][2]);
        q[0] = 0.25f * s;
        q[1] = (m->a[1][0] + m->a[0][1]) / s;
        q[2] = (m->a[0][2] + m->a[2][0]) / s;
        q[3] = (m->a[2][1] - m->a[1][2]) / s;
    } else if (m->a[1][1] > m->a[2][2]) {
        const float s = 2.0f * sqrtf(1.0f - m->a[0][0] + m->a[1][1] - m->a[2][2]);
        q[0] = (m->a[1][0] + m->a[0][1]) / s;
        q[1] = 0.25f * s;
        q[2] = (m->a[2][1] + m->a[1][2]) / s;
        q[3] = (m->a[0][2] - m->a[2][0]) / s;
    } else {
        const float s = 2.0f * sqrtf(1.0f - m->a[0][0] - m->a[1][1] + m->a[2][2]);
        q[0] = (m->a[0][2] + m->a[2][0]) / s;
        q[1] = (m->a[2][1] + m->a[1][2]) / s;
        q[2] = 0.25f * s;
        q[3] = (m->a[1][0] - m->a[0][1]) / s;
    }
    const float len = sqrtf(((q[0] * q[0] + q[1] * q[1]) + q[2] * q[2]) + q[3] * q[3]);
    for (int k = 0; k < 4; ++k) q[k] = q[k] / len;
}

/* lodFinalize (lod_finalize.slang:43-97): moments -> one packed Gaussian a group. */
void or_lod_finalize(const float* moments, uint32_t groups, uint32_t keep, uint32_t sh_words, float* positions,
                     uint32_t* shape, uint32_t* sh) {
    const uint32_t stride = 13u + keep * 3u;
    for (uint32_t g = 0; g < groups; ++g) {
        const float* m = moments + (size_t)g * stride;
        const float weight = fmax_(m[0], 1e-30f);
        const float mu[3] = {m[1] / weight, m[2] / weight, m[3] / weight};
        m3 cov = {{{m[4] / weight, m[5] / weight, m[6] / weight},
                   {m[5] / weight, m[7] / weight, m[8] / weight},
                   {m[6] / weight, m[8] / weight, m[9] / weight}}};
        for (int i = 0; i < 3; ++i)
            for (int k = 0; k < 3; ++k) cov.a[i][k] = cov.a[i][k] - mu[i] * mu[k];
        m3 axes;
        jacobi3(&cov, &axes);
        const float det = axes.a[0][0] * (axes.a[1][1] * axes.a[2][2] - axes.a[1][2] * axes.a[2][1]) -
                          axes.a[0][1] * (axes.a[1][0] * axes.a[2][2] - axes.a[1][2] * axes.a[2][0]) +
                          axes.a[0][2] * (axes.a[1][0] * axes.a[2][1] - axes.a[1][1] * axes.a[2][0]);
        if (det < 0.0f) {
            axes.a[0][2] = -axes.a[0][2];
            axes.a[1][2] = -axes.a[1][2];
            axes.a[2][2] = -axes.a[2][2];
        }
        const float s[3] = {sqrtf(fmax_(cov.a[0][0], 1e-14f)), sqrtf(fmax_(cov.a[1][1], 1e-14f)),
                            sqrtf(fmax_(cov.a[2][2], 1e-14f))};
        const float smallest = fmin_(fmin_(s[0], s[1]), s[2]);
        const float area = s[0] * s[1] * s[2] / fmax_(smallest, 1e-20f);
        const float opacity = fmin_(m[0] / fmax_(area, 1e-30f), 0.99f);
        float* pos = positions + (size_t)g * 4;
        pos[0] = mu[0]; pos[1] = mu[1]; pos[2] = mu[2]; pos[3] = opacity;
        float q[4];
        quaternion_of_axes(&axes, q);
        uint32_t* row = shape + (size_t)g * 4;
        row[0] = or_encode_quaternion(q);
        row[1] = or_pack_halves(half_safe(or_logf(s[0])), half_safe(or_logf(s[1])));
        row[2] = or_pack_halves(half_safe(or_logf(s[2])), half_safe(m[10] / weight));
        row[3] = or_pack_halves(half_safe(m[11] / weight), half_safe(m[12] / weight));
        if (keep == 0) continue;
        uint32_t word = 0, filled = 0, written = 0;
        for (uint32_t h = 0; h < keep * 3; ++h) {
            word |= (uint32_t)or_float_to_half(half_safe(m[13 + h] / weight)) << (filled * 16);
            if (++filled == 2) {
                sh[(size_t)g * sh_words + written++] = word;
                word = 0;
                filled = 0;
            }
        }
        if (filled == 1) sh[(size_t)g * sh_words + written] = word;
    }
}

/* projectedEdge (lod_cut.slang:13-24): the pixels a level-`level` cell spans
 * at its nearest point to the eye. */
float or_lod_projected_edge(const or_cut_params* p, uint32_t cell, uint32_t level) {
    const float edge = p->extent / (float)(1u << level);
    if (p->orthographic) return edge * p->pixels_per_unit;
    const float lo[3] = {p->lo[0] + (float)compact_bits(cell >> 2) * edge, p->lo[1] + (float)compact_bits(cell >> 1) * edge,
                         p->lo[2] + (float)compact_bits(cell) * edge};
    float d[3];
    for (int k = 0; k < 3; ++k) d[k] = clampf_(p->eye[k], lo[k], lo[k] + edge) - p->eye[k];
    const float distance = fmax_(sqrtf((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]), 1e-6f);
    return edge * p->pixels_per_unit / distance;
}

static int lod_small(const or_cut_params* p, uint32_t cell, uint32_t level) {
    return p->threshold > 0.0f && or_lod_projected_edge(p, cell, level) <= p->threshold;
}

/* lodCutGroups, and lodCutFinest with every chunk resident (lod_cut.slang:
 * 30-80): drawn when the cell is small and (below the coarsest stored level)
 * its parent's is not. */
void or_lod_cut_groups(const or_cut_params* p, const uint32_t* cells, uint32_t count, uint32_t level, int coarsest,
                       uint32_t* selected) {
    for (uint32_t g = 0; g < count; ++g) {
        const int here = lod_small(p, cells[g], level);
        const int above = !coarsest && lod_small(p, cells[g] >> 3, level - 1);
        selected[g] = (here && !above) ? 1u : 0u;
    }
}

/* lodCutSplats through lodCutFinest's state (lod_cut.slang:56-92): a splat is
 * drawn when its finest-level cell is not small. */
void or_lod_cut_splats(const or_cut_params* p, const uint32_t* groups, uint32_t n, const uint32_t* finest_cells,
                       uint32_t finest_level, uint32_t* selected) {
    for (uint32_t i = 0; i < n; ++i) selected[i] = lod_small(p, finest_cells[groups[i]], finest_level) ? 0u : 1u;
}
