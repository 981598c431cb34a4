// Device arithmetic of the splat path: exp/log, fp16 and quaternion decode,
// the footprint cut and the exact tile test, harmonics and the SplatEdit.
//
// Every function here feeds the tile counts, the depth keys or the projected
// record, all of which must agree bit for bit with the CPU restatement of the
// reference. So: no FMA contraction (the library is built with
// -ffp-contract=off), IEEE division and sqrt (-fhip-fp32-correctly-rounded-
// divide-sqrt), operations in the reference's source order, and exp/log by a
// fixed published method instead of the device library's, whose last bits
// differ from any host library's.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "frame.h"

namespace lrt::dev {

__device__ __forceinline__ float bitsToFloat(uint32_t u) { return __uint_as_float(u); }
__device__ __forceinline__ uint32_t floatToBits(float f) { return __float_as_uint(f); }
__device__ __forceinline__ float fminS(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float fmaxS(float a, float b) { return a > b ? a : b; }
__device__ __forceinline__ float clampS(float v, float lo, float hi) { return fminS(fmaxS(v, lo), hi); }

// exp(x): Cody & Waite reduction x = k ln2 + r with ln2 split in a 15-bit head
// (so k * head is exact for |k| < 256) and a tail, then the degree-7 Taylor
// polynomial of e^r in Horner form (|r| <= ln2/2, truncation < 4e-9 relative),
// then an exact scaling by 2^k (two steps when the result is subnormal).
__device__ __forceinline__ float expF(float x) {
    if (x != x) return x;
    if (x > 88.72283935546875f) return __int_as_float(0x7F800000);
    if (x < -103.97208404541016f) return 0.0f;
    const float kf = __builtin_rintf(x * 1.4426950408889634f);
    int k = static_cast<int>(kf);
    float r = x - kf * 0.693145751953125f;
    r = r - kf * 1.428606765330187e-06f;
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
    if (k >= -126) return p * bitsToFloat(static_cast<uint32_t>(k + 127) << 23);
    return (p * bitsToFloat(static_cast<uint32_t>(k + 64 + 127) << 23)) * bitsToFloat(static_cast<uint32_t>(63) << 23);
}

// log(x) = e ln2 + log(m), m in [sqrt(1/2), sqrt(2)]; log(m) = 2 atanh(s) with
// s = f/(2+f), f = m - 1, written as f - s (f - R(s^2)) where R is the atanh
// series' tail 2z/3 + 2z^2/5 + ... + 2z^5/11.
__device__ __forceinline__ float logF(float x) {
    if (x != x) return x;
    if (x < 0.0f) return __int_as_float(0x7FC00000);
    if (x == 0.0f) return __int_as_float(0xFF800000);
    uint32_t b = floatToBits(x);
    if (b == 0x7F800000u) return x;
    int e = 0;
    if (b < 0x00800000u) {
        x = x * 33554432.0f;
        b = floatToBits(x);
        e = -25;
    }
    e += static_cast<int>(b >> 23) - 127;
    float m = bitsToFloat((b & 0x007FFFFFu) | 0x3F800000u);
    if (m > 1.41421353816986083984375f) {
        m = m * 0.5f;
        e += 1;
    }
    const float f = m - 1.0f;
    const float s = f / (2.0f + f);
    const float z = s * s;
    const float r = z * (2.0f / 3.0f + z * (2.0f / 5.0f + z * (2.0f / 7.0f + z * (2.0f / 9.0f + z * (2.0f / 11.0f)))));
    const float lm = f - s * (f - r);
    const float ef = static_cast<float>(e);
    return ef * 0.693145751953125f + (ef * 1.428606765330187e-06f + lm);
}

// fp16 -> fp32, exact, a NaN keeping its payload bits (sign, all-ones
// exponent, mantissa << 13). The hardware conversion (v_cvt_f32_f16) gives
// exactly that for every half but a signalling NaN, which it quiets (sets
// bit 22); for a NaN, bit 22 is put back from the half's top mantissa bit.
// Branch-free.
__device__ __forceinline__ float halfToFloat(uint32_t h) {
    const float hw = static_cast<float>(__builtin_bit_cast(_Float16, static_cast<uint16_t>(h)));
    const uint32_t payload = (floatToBits(hw) & ~0x00400000u) | ((h & 0x200u) << 13);
    return hw != hw ? bitsToFloat(payload) : hw;
}
__device__ __forceinline__ float lowHalf(uint32_t w) { return halfToFloat(w & 0xFFFFu); }
__device__ __forceinline__ float highHalf(uint32_t w) { return halfToFloat(w >> 16); }

// s / 1023 correctly rounded, for an integer s in [0, 1023]: the product with
// the rounded reciprocal, corrected once by FMA, equals the division for each
// of the 1,024 values (checked exhaustively in exact arithmetic). Three
// instructions instead of a division's ten.
__device__ __forceinline__ float div1023(float s) {
    constexpr float kR = 1.0f / 1023.0f;
    const float q = s * kR;
    return __builtin_fmaf(__builtin_fmaf(-q, 1023.0f, s), kR, q);
}

// Smallest-three quaternion (reference packing.slang:52-69): three 10-bit
// components in [-sqrt(1/2), sqrt(1/2)], the largest from unit length.
__device__ __forceinline__ void decodeQuaternion(uint32_t word, float q[4]) {
    const uint32_t largest = word & 3u;
    const float kSqrtHalf = 0.70710678f;
    float s0 = static_cast<float>((word >> 22) & 1023u);
    float s1 = static_cast<float>((word >> 12) & 1023u);
    float s2 = static_cast<float>((word >> 2) & 1023u);
    s0 = (div1023(s0) * 2.0f - 1.0f) * kSqrtHalf;
    s1 = (div1023(s1) * 2.0f - 1.0f) * kSqrtHalf;
    s2 = (div1023(s2) * 2.0f - 1.0f) * kSqrtHalf;
    const float rest = __builtin_sqrtf(fmaxS(0.0f, 1.0f - ((s0 * s0 + s1 * s1) + s2 * s2)));
    // Place the three stored components around the largest one.
    q[0] = largest == 0 ? rest : s0;
    q[1] = largest == 1 ? rest : (largest == 0 ? s0 : s1);
    q[2] = largest == 2 ? rest : (largest <= 1 ? s1 : s2);
    q[3] = largest == 3 ? rest : s2;
}

// The squared Mahalanobis radius a splat is cut at (frame.slang:105-107).
__device__ __forceinline__ float cutoffPower(float alpha) {
    return fminS(2.0f * logF(fmaxS(255.0f * alpha, 1.0001f)), 9.0f);
}

// Exact minimum of a dx^2 + 2b dx dy + c dy^2 over a box (frame.slang:113-130).
__device__ __forceinline__ float minPowerOverBox(float a, float b, float c, float lox, float loy, float hix, float hiy) {
    if (lox <= 0.0f && hix >= 0.0f && loy <= 0.0f && hiy >= 0.0f) return 0.0f;
    float best = 3.4e38f;
#pragma unroll
    for (int e = 0; e < 2; ++e) {
        const float x = e == 0 ? lox : hix;
        const float y = c > 1e-12f ? clampS(-b * x / c, loy, hiy) : loy;
        best = fminS(best, a * x * x + 2.0f * b * x * y + c * y * y);
    }
#pragma unroll
    for (int e = 0; e < 2; ++e) {
        const float y = e == 0 ? loy : hiy;
        const float x = a > 1e-12f ? clampS(-b * y / a, lox, hix) : lox;
        best = fminS(best, a * x * x + 2.0f * b * x * y + c * y * y);
    }
    return best;
}

// Can the Gaussian (power a x^2 + 2 b x y + c y^2 about its centre) reach
// cutoff at some point of the box [lox, hix] x [loy, hiy] (coordinates about
// the centre)? The power is convex, so its minimum over the box is 0 when the
// centre is inside and otherwise lies on an edge, at the clamped minimiser of
// the edge's 1-D quadratic. The answer may be yes for a box that no pixel
// reaches, never no for one that a pixel reaches: the slack bounds the float
// rounding of this evaluation and of the per-pixel one, both below a few ulps
// of (a + 2|b| + c) * r2 with r2 the largest squared coordinate in the box,
// ten times over; the edge minimisers use the hardware reciprocal, whose error
// moves the value only to second order. A degenerate form is always kept.
__device__ __forceinline__ bool quadrantReach(float a, float b, float c, float cutoff, float lox, float loy,
                                              float hix, float hiy) {
    const bool centreInside = lox <= 0.0f && hix >= 0.0f && loy <= 0.0f && hiy >= 0.0f;
    const bool degenerate = !(a > 1e-12f && c > 1e-12f);
    const float ra = __builtin_amdgcn_rcpf(a), rc = __builtin_amdgcn_rcpf(c);
    const auto form = [&](float x, float y) { return __builtin_fmaf(a * x, x, __builtin_fmaf(2.0f * b * x, y, c * y * y)); };
    float best = form(lox, fminS(fmaxS(-b * lox * rc, loy), hiy));
    best = fminS(best, form(hix, fminS(fmaxS(-b * hix * rc, loy), hiy)));
    best = fminS(best, form(fminS(fmaxS(-b * loy * ra, lox), hix), loy));
    best = fminS(best, form(fminS(fmaxS(-b * hiy * ra, lox), hix), hiy));
    const float r2 = fmaxS(fmaxS(lox * lox, hix * hix), fmaxS(loy * loy, hiy * hiy));
    const float slack = 4e-6f * (a + 2.0f * __builtin_fabsf(b) + c) * r2 + 1e-5f * cutoff + 1e-6f;
    return centreInside || degenerate || best <= cutoff + slack;
}

// Does tile (tx, ty), 1 << tileShift pixels square, hold a pixel centre inside
// the footprint (frame.slang:133-138, whose tiles are 16 pixels: tileShift 4)?
__device__ __forceinline__ bool tileTouched(uint32_t width, uint32_t height, float cx, float cy, float a, float b,
                                            float c, float cutoff, uint32_t tx, uint32_t ty, uint32_t tileShift) {
    const float fx = static_cast<float>(tx << tileShift) + 0.5f;
    const float fy = static_cast<float>(ty << tileShift) + 0.5f;
    const uint32_t ex = (tx + 1u) << tileShift < width ? (tx + 1u) << tileShift : width;
    const uint32_t ey = (ty + 1u) << tileShift < height ? (ty + 1u) << tileShift : height;
    const float lx = static_cast<float>(ex - 1u) + 0.5f;
    const float ly = static_cast<float>(ey - 1u) + 0.5f;
    return minPowerOverBox(a, b, c, fx - cx, fy - cy, lx - cx, ly - cy) <= cutoff;
}

// The tile rectangle, inclusive, of a footprint with half-widths ex, ey about
// (cx, cy) (frame.slang:149-161), in tiles of 1 << tileShift pixels; empty
// (x0 > x1) when it is off screen.
__device__ __forceinline__ void tileRectOfExtent(uint32_t width, uint32_t height, float cx, float cy, float ex,
                                                 float ey, uint32_t rect[4], uint32_t tileShift) {
    const float x1 = cx - ex, y1 = cy - ey, x2 = cx + ex, y2 = cy + ey;
    if (x2 < 0.0f || y2 < 0.0f || x1 >= static_cast<float>(width) || y1 >= static_cast<float>(height)) {
        rect[0] = 1; rect[1] = 1; rect[2] = 0; rect[3] = 0;
        return;
    }
    rect[0] = static_cast<uint32_t>(fmaxS(x1, 0.0f)) >> tileShift;
    rect[1] = static_cast<uint32_t>(fmaxS(y1, 0.0f)) >> tileShift;
    rect[2] = static_cast<uint32_t>(fminS(x2, static_cast<float>(width) - 1.0f)) >> tileShift;
    rect[3] = static_cast<uint32_t>(fminS(y2, static_cast<float>(height) - 1.0f)) >> tileShift;
}

// The tile rectangle of a Gaussian's footprint: its extent from the 2D
// covariance, sqrt(cutoff sigma_xx) and sqrt(cutoff sigma_yy) (frame.slang:142-146).
__device__ __forceinline__ void tileRect(uint32_t width, uint32_t height, float cx, float cy, float a, float b,
                                         float c, float cutoff, uint32_t rect[4], uint32_t tileShift) {
    const float det = a * c - b * b;
    const float inv = det > 1e-20f ? 1.0f / det : 0.0f;
    const float ex = __builtin_sqrtf(fmaxS(cutoff * c * inv, 0.0f));
    const float ey = __builtin_sqrtf(fmaxS(cutoff * a * inv, 0.0f));
    tileRectOfExtent(width, height, cx, cy, ex, ey, rect, tileShift);
}

__device__ __forceinline__ float shflF(float v, uint32_t lane) { return __shfl(v, static_cast<int>(lane), 64); }
__device__ __forceinline__ uint32_t shflU(uint32_t v, uint32_t lane) {
    return static_cast<uint32_t>(__shfl(static_cast<int>(v), static_cast<int>(lane), 64));
}

__device__ __forceinline__ uint32_t waveMin(uint32_t v) {
#pragma unroll
    for (uint32_t d = 1; d < 64; d <<= 1) v = min(v, static_cast<uint32_t>(__shfl_xor(static_cast<int>(v), d, 64)));
    return v;
}
__device__ __forceinline__ uint32_t waveMax(uint32_t v) {
#pragma unroll
    for (uint32_t d = 1; d < 64; d <<= 1) v = max(v, static_cast<uint32_t>(__shfl_xor(static_cast<int>(v), d, 64)));
    return v;
}

// The wave's part of its workgroup's depth-key range: the least and greatest
// depth key (bits(tz) >> 7) of its drawn lanes, written by lane 0 to
// sWaves[2 w] and sWaves[2 w + 1] (w the wave in the workgroup). Every lane
// of the wave must take part.
__device__ __forceinline__ void waveDepthRange(bool drawn, float tz, uint32_t* sWaves) {
    const uint32_t key = (floatToBits(tz) & 0x7FFFFFFFu) >> 7;
    const uint32_t lo = waveMin(drawn ? key : 0xFFFFFFFFu), hi = waveMax(drawn ? key : 0u);
    if (__lane_id() == 0) {
        sWaves[2 * (threadIdx.x >> 6)] = lo;
        sWaves[2 * (threadIdx.x >> 6) + 1] = hi;
    }
}

// After a barrier that follows every wave's waveDepthRange: thread 0 writes
// the workgroup's entry (kProjectBlock / 64 waves folded).
__device__ __forceinline__ void storeDepthRange(const DepthBlocks& depth, const uint32_t* sWaves) {
    if (depth.entries == nullptr || threadIdx.x != 0) return;
    uint32_t lo = 0xFFFFFFFFu, hi = 0u;
#pragma unroll
    for (uint32_t w = 0; w < kProjectBlock / 64; ++w) {
        lo = min(lo, sWaves[2 * w]);
        hi = max(hi, sWaves[2 * w + 1]);
    }
    uint32_t* e = depth.entries + 2ull * (depth.first + blockIdx.x);
    e[0] = lo;
    e[1] = hi;
}

// The wave's candidate tiles, flattened: lane L's rectangle (area[L] tiles in
// row-major order) is candidates [start[L], start[L] + area[L]) of one list,
// start the exclusive sum of the areas over the lanes. An iteration tests 64
// consecutive candidates, one a lane, so a wave walks ceil(total / 64) rounds
// however its areas are spread, instead of its largest area. Every lane of
// the wave must take part (a lane with nothing to count brings area 0).
struct WaveTiles {
    uint32_t incl, start, area, total;
};

__device__ __forceinline__ WaveTiles waveTiles(uint32_t area) {
    const uint32_t lane = __lane_id();
    uint32_t incl = area;
#pragma unroll
    for (uint32_t d = 1; d < 64; d <<= 1) {
        const uint32_t o = static_cast<uint32_t>(__shfl_up(static_cast<int>(incl), d, 64));
        incl += lane >= d ? o : 0u;
    }
    return WaveTiles{incl, incl - area, area, shflU(incl, 63)};
}

// The lane whose run holds candidate `cand`: the first whose inclusive end
// passes it (a binary search over the monotone ends).
__device__ __forceinline__ uint32_t ownerOf(uint32_t incl, uint32_t cand) {
    uint32_t lo = 0;
#pragma unroll
    for (uint32_t step = 32; step > 0; step >>= 1) lo += shflU(incl, lo + step - 1u) <= cand ? step : 0u;
    return lo;
}

// The bits of this lane's run that fall in the iteration from candidate `base`.
__device__ __forceinline__ uint64_t runMask(uint32_t start, uint32_t area, uint32_t base) {
    const int64_t lo = static_cast<int64_t>(start) - base, hi = lo + area;
    const int64_t l = lo < 0 ? 0 : (lo > 64 ? 64 : lo), h = hi < 0 ? 0 : (hi > 64 ? 64 : hi);
    if (l >= h) return 0ull;
    const uint64_t upto = h == 64 ? ~0ull : (1ull << h) - 1ull;
    return upto & ~((1ull << l) - 1ull);
}

// k / w and k % w for a candidate's place k in a rectangle w tiles wide: below
// 2^20 a float estimate (relative error under 2^-22, so off by at most one),
// corrected; exact division past that.
__device__ __forceinline__ void rowColumn(uint32_t k, uint32_t w, bool exact, uint32_t& row, uint32_t& col) {
    if (exact) {
        row = k / w;
        col = k - row * w;
        return;
    }
    int q = static_cast<int>(static_cast<float>(k) * __builtin_amdgcn_rcpf(static_cast<float>(w)));
    int r = static_cast<int>(k) - q * static_cast<int>(w);
    if (r < 0) {
        q -= 1;
        r += static_cast<int>(w);
    } else if (r >= static_cast<int>(w)) {
        q += 1;
        r -= static_cast<int>(w);
    }
    row = static_cast<uint32_t>(q);
    col = static_cast<uint32_t>(r);
}

// The number of tiles of each lane's rect (inclusive x0, y0, x1, y1) that its
// footprint touches (tileTouched), every lane's rectangle among the wave's
// flattened candidates. `want` false: the lane counts nothing (its rect is not
// read). Tiles are 1 << tileShift pixels square. Every lane of the wave must
// reach the call.
__device__ __forceinline__ uint32_t countTouchedTilesWave(uint32_t width, uint32_t height, float cx, float cy,
                                                          float a, float b, float c, float cutoff,
                                                          const uint32_t rect[4], bool want, uint32_t tileShift) {
    const uint32_t lane = __lane_id();
    const uint32_t w = want ? rect[2] - rect[0] + 1u : 1u;
    const WaveTiles wt = waveTiles(want ? w * (rect[3] - rect[1] + 1u) : 0u);
    const bool exact = wt.total >= (1u << 20);
    uint32_t touched = 0;
    for (uint32_t base = 0; base < wt.total; base += 64) {
        const uint32_t cand = base + lane;
        const uint32_t o = ownerOf(wt.incl, cand);
        const uint32_t oStart = shflU(wt.start, o), oW = shflU(w, o);
        const uint32_t ox0 = shflU(want ? rect[0] : 0u, o), oy0 = shflU(want ? rect[1] : 0u, o);
        const float ocx = shflF(cx, o), ocy = shflF(cy, o), oa = shflF(a, o), ob = shflF(b, o), oc = shflF(c, o);
        const float ocut = shflF(cutoff, o);
        bool hit = false;
        if (cand < wt.total) {
            uint32_t row, col;
            rowColumn(cand - oStart, oW, exact, row, col);
            hit = tileTouched(width, height, ocx, ocy, oa, ob, oc, ocut, ox0 + col, oy0 + row, tileShift);
        }
        touched += static_cast<uint32_t>(__popcll(__ballot(hit) & runMask(wt.start, wt.area, base)));
    }
    return touched;
}

__device__ __forceinline__ uint32_t degreeOf(uint32_t rest) {
    return rest >= 15 ? 3 : rest >= 8 ? 2 : rest >= 3 ? 1 : 0;
}

// The view-dependent part of the colour (reference sh.slang:25-50), the
// coefficients already unpacked to fp32, rgb per basis: k[basis * 3 + channel].
__device__ __forceinline__ void evaluateRest(const float* k, uint32_t degree, float x, float y, float z, float out[3]) {
    const float kC1 = 0.4886025119029199f;
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) out[ch] = 0.0f;
    if (degree == 0) return;
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) out[ch] = -kC1 * y * k[0 + ch] + kC1 * z * k[3 + ch] - kC1 * x * k[6 + ch];
    if (degree < 2) return;
    const float xx = x * x, yy = y * y, zz = z * z;
    const float xy = x * y, yz = y * z, xz = x * z;
    const float c20 = 1.0925484305920792f, c21 = -1.0925484305920792f, c22 = 0.31539156525252005f,
                c23 = -1.0925484305920792f, c24 = 0.5462742152960396f;
#pragma unroll
    for (int ch = 0; ch < 3; ++ch)
        out[ch] = out[ch] + (c20 * xy * k[9 + ch] + c21 * yz * k[12 + ch] + c22 * (2.0f * zz - xx - yy) * k[15 + ch] +
                             c23 * xz * k[18 + ch] + c24 * (xx - yy) * k[21 + ch]);
    if (degree < 3) return;
    const float c30 = -0.5900435899266435f, c31 = 2.890611442640554f, c32 = -0.4570457994644658f,
                c33 = 0.3731763325901154f, c34 = -0.4570457994644658f, c35 = 1.445305721320277f,
                c36 = -0.5900435899266435f;
#pragma unroll
    for (int ch = 0; ch < 3; ++ch)
        out[ch] = out[ch] + (c30 * y * (3.0f * xx - yy) * k[24 + ch] + c31 * xy * z * k[27 + ch] +
                             c32 * y * (4.0f * zz - xx - yy) * k[30 + ch] +
                             c33 * z * (2.0f * zz - 3.0f * xx - 3.0f * yy) * k[33 + ch] +
                             c34 * x * (4.0f * zz - xx - yy) * k[36 + ch] + c35 * z * (xx - yy) * k[39 + ch] +
                             c36 * x * (xx - 3.0f * yy) * k[42 + ch]);
}

struct EditFate {
    bool removed;
    bool graded;
    float opacity;
};

// What a SplatEdit does to a splat (reference edit.slang:46-78).
__device__ __forceinline__ EditFate editFate(const SplatEditParams& e, float px, float py, float pz, float opacity,
                                             float sx, float sy, float sz) {
    EditFate fate{false, false, opacity};
    if (e.active == 0) return fate;
    if (e.minOpacity > 0.0f && opacity < e.minOpacity) {
        fate.removed = true;
        return fate;
    }
    if (e.maxScale > 0.0f && fmaxS(fmaxS(__builtin_fabsf(sx), __builtin_fabsf(sy)), __builtin_fabsf(sz)) > e.maxScale) {
        fate.removed = true;
        return fate;
    }
    const float rx = px - e.centre[0], ry = py - e.centre[1], rz = pz - e.centre[2];
    bool inside = e.shape == 1 ? ((rx * rx + ry * ry) + rz * rz) <= e.size[0] * e.size[0]
                               : (__builtin_fabsf(rx) <= e.size[0] && __builtin_fabsf(ry) <= e.size[1] &&
                                  __builtin_fabsf(rz) <= e.size[2]);
    if (e.invert != 0) inside = !inside;
    if ((e.mode == 0 && !inside) || (e.mode == 1 && inside)) {
        fate.removed = true;
        return fate;
    }
    if (inside) {
        fate.graded = true;
        fate.opacity = opacity * e.opacity;
        fate.removed = !(fate.opacity >= 1.0f / 255.0f);
    }
    return fate;
}

// The grade (reference edit.slang:81-85).
__device__ __forceinline__ void editGrade(const SplatEditParams& e, float c[3]) {
    const float r = c[0] * e.tint[0] * e.brightness;
    const float g = c[1] * e.tint[1] * e.brightness;
    const float b = c[2] * e.tint[2] * e.brightness;
    const float luma = (r * 0.2126f + g * 0.7152f) + b * 0.0722f;
    c[0] = fmaxS(luma + (r - luma) * e.saturation, 0.0f);
    c[1] = fmaxS(luma + (g - luma) * e.saturation, 0.0f);
    c[2] = fmaxS(luma + (b - luma) * e.saturation, 0.0f);
}

// The reference's sRGB curve, extended past 1 (color.slang:7-10), with pow as
// exp(2.4 log y) by the functions above.
__device__ __forceinline__ float srgbToLinear(float c) {
    c = fmaxS(c, 0.0f);
    return c <= 0.04045f ? c / 12.92f : expF(2.4f * logF((c + 0.055f) / 1.055f));
}

// The view of a batch that holds batch tile g: the last view whose tiles
// start at or before it (views in ascending tileBase).
__device__ __forceinline__ uint32_t viewOfTile(const ViewInfo* __restrict__ views, uint32_t viewCount, uint32_t g) {
    uint32_t lo = 0, hi = viewCount;   // views[lo].tileBase <= g < views[hi].tileBase
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (views[mid].tileBase <= g) lo = mid; else hi = mid;
    }
    return lo;
}

}   // namespace lrt::dev
