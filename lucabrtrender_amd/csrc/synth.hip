// Host-side packing of splats into the GpuSplats layout, and the seeded
// synthetic scenes the rasteriser is measured and checked on.
//
// Packing follows the reference's splat_decode (shaders/lrt/scene/
// splat_decode.slang:17-87) through packing.h, the decode kernel's own.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/lrt_synth.h"
#include "packing.h"
#include "synth.h"

namespace lrt::synth {

void packSplat(const SplatFields& s, uint32_t restPerColour, uint32_t shWords, float* position, uint32_t* shape,
               uint32_t* sh) {
    position[0] = s.position[0];
    position[1] = s.position[1];
    position[2] = s.position[2];
    position[3] = std::min(std::max(s.opacity, 0.0f), 1.0f);
    float q[4] = {s.rotation[0], s.rotation[1], s.rotation[2], s.rotation[3]};
    const float ql = std::sqrt(((q[0] * q[0] + q[1] * q[1]) + q[2] * q[2]) + q[3] * q[3]);
    if (!(ql > 1e-8f)) {
        q[0] = 0.0f; q[1] = 0.0f; q[2] = 0.0f; q[3] = 1.0f;
    }
    shape[0] = pack::encodeQuaternion(q);
    shape[1] = pack::packHalves(s.logScale[0], s.logScale[1]);
    shape[2] = pack::packHalves(s.logScale[2], s.base[0]);
    shape[3] = pack::packHalves(s.base[1], s.base[2]);
    if (sh == nullptr) return;
    for (uint32_t w = 0; w < shWords; ++w) sh[w] = 0;
    const uint32_t halves = restPerColour * 3;
    for (uint32_t k = 0; k < halves; ++k) {
        sh[k / 2] |= static_cast<uint32_t>(pack::floatToHalf(pack::halfSafe(s.rest[k]))) << ((k & 1u) * 16);
    }
}

// splitmix64, the per-splat stream seed: every splat's draws are independent of
// every other's, so a scene is generated in parallel and any prefix of it is the
// same scene cut short.
uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// The reference's test LCG (tests/render/SplatFixtures.h:48-55): 24-bit uniforms.
struct Lcg {
    uint64_t state;
    float next() {
        state = state * 6364136223846793005ull + 1442695040888963407ull;
        return static_cast<float>((state >> 40) & 0xFFFFFF) / 16777216.0f;
    }
    float range(float lo, float hi) { return lo + (hi - lo) * next(); }
    // Box-Muller from two 24-bit uniforms, the first kept off zero.
    float normal(float mean, float sd) {
        state = state * 6364136223846793005ull + 1442695040888963407ull;
        const double u1 = (static_cast<double>((state >> 40) & 0xFFFFFF) + 0.5) / 16777216.0;
        state = state * 6364136223846793005ull + 1442695040888963407ull;
        const double u2 = static_cast<double>((state >> 40) & 0xFFFFFF) / 16777216.0;
        const double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
        return static_cast<float>(mean + sd * z);
    }
};

void drawSplat(const SceneSpec& spec, uint32_t i, const float* clusterCentres, const float* clusterSd, SplatFields& out) {
    Lcg rng{splitmix64(spec.seed * 0x100000001B3ull + i) ^ 0x853c49e6748fea9bull};
    const bool near = spec.kind == SceneKind::Mixed && rng.next() < spec.nearFraction;
    if (spec.kind == SceneKind::Clusters) {
        const uint32_t c = std::min(static_cast<uint32_t>(rng.next() * static_cast<float>(spec.clusters)), spec.clusters - 1);
        for (int k = 0; k < 3; ++k) out.position[k] = rng.normal(clusterCentres[c * 3 + k], clusterSd[c * 3 + k]);
    } else {
        for (int k = 0; k < 3; ++k) out.position[k] = rng.range(-1.0f, 1.0f) + spec.offset[k];
        if (near) {
            // Pulled toward the camera: the same cube scaled into [0.2, 1] of its distance.
            const float t = rng.range(0.2f, 1.0f);
            for (int k = 0; k < 3; ++k) out.position[k] = out.position[k] * t;
        }
    }
    const float lsMean = near ? spec.nearLogScaleMean : spec.logScaleMean;
    const float lsSd = near ? spec.nearLogScaleSd : spec.logScaleSd;
    for (int k = 0; k < 3; ++k) out.logScale[k] = rng.normal(lsMean, lsSd);
    for (int k = 0; k < 4; ++k) out.rotation[k] = rng.normal(0.0f, 1.0f);
    const float logit = rng.normal(spec.opacityLogitMean, spec.opacityLogitSd);
    out.opacity = static_cast<float>(1.0 / (1.0 + std::exp(-static_cast<double>(logit))));
    for (int k = 0; k < 3; ++k) out.base[k] = rng.next();
    const uint32_t bases = spec.degree == 0 ? 0u : spec.degree == 1 ? 3u : spec.degree == 2 ? 8u : 15u;
    for (uint32_t b = 0; b < bases; ++b) {
        const float band = b < 3 ? 1.0f : b < 8 ? 2.0f : 3.0f;
        for (int ch = 0; ch < 3; ++ch) out.rest[b * 3 + static_cast<uint32_t>(ch)] = rng.normal(0.0f, spec.shSd / band);
    }
}

void generate(const SceneSpec& spec, uint32_t count, float* positions, uint32_t* shape, uint32_t* sh) {
    const uint32_t rest = spec.degree == 0 ? 0u : spec.degree == 1 ? 3u : spec.degree == 2 ? 8u : 15u;
    const uint32_t words = rest == 0 ? 1u : (rest * 3 + 1) / 2;
    std::vector<float> centres(spec.clusters * 3), sds(spec.clusters * 3);
    if (spec.kind == SceneKind::Clusters) {
        Lcg rng{splitmix64(spec.seed ^ 0xC1u) ^ 0x853c49e6748fea9bull};
        for (uint32_t c = 0; c < spec.clusters; ++c) {
            for (int k = 0; k < 3; ++k) centres[c * 3 + k] = rng.range(-spec.boxHalf, spec.boxHalf);
            for (int k = 0; k < 3; ++k) sds[c * 3 + k] = rng.range(spec.clusterSdLo, spec.clusterSdHi);
        }
    }
    const uint32_t threads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const auto work = [&](uint32_t t) {
        SplatFields s;
        const uint64_t lo = uint64_t{count} * t / threads, hi = uint64_t{count} * (t + 1) / threads;
        for (uint64_t i = lo; i < hi; ++i) {
            drawSplat(spec, static_cast<uint32_t>(i), centres.data(), sds.data(), s);
            packSplat(s, rest, words, positions + i * 4, shape + i * 4, sh != nullptr ? sh + i * words : nullptr);
        }
    };
    std::vector<std::thread> pool;
    for (uint32_t t = 1; t < threads; ++t) pool.emplace_back(work, t);
    work(0);
    for (std::thread& th : pool) th.join();
}

}   // namespace lrt::synth

extern "C" {

lrt_status lrt_synth_scene(const lrt_scene_spec* spec, uint32_t count, float* positions, uint32_t* shape,
                           uint32_t* sh) {
    if (spec == nullptr || positions == nullptr || shape == nullptr) return LRT_INVALID_ARGUMENT;
    if (spec->kind > 2 || spec->degree > 3 || (spec->kind == 1 && spec->clusters == 0)) return LRT_INVALID_ARGUMENT;
    if (spec->degree > 0 && sh == nullptr) return LRT_INVALID_ARGUMENT;
    lrt::synth::SceneSpec s;
    s.kind = static_cast<lrt::synth::SceneKind>(spec->kind);
    s.seed = spec->seed;
    s.degree = spec->degree;
    for (int k = 0; k < 3; ++k) s.offset[k] = spec->offset[k];
    s.logScaleMean = spec->log_scale_mean;
    s.logScaleSd = spec->log_scale_sd;
    s.opacityLogitMean = spec->opacity_logit_mean;
    s.opacityLogitSd = spec->opacity_logit_sd;
    s.shSd = spec->sh_sd;
    s.clusters = spec->clusters;
    s.boxHalf = spec->box_half;
    s.clusterSdLo = spec->cluster_sd_lo;
    s.clusterSdHi = spec->cluster_sd_hi;
    s.nearFraction = spec->near_fraction;
    s.nearLogScaleMean = spec->near_log_scale_mean;
    s.nearLogScaleSd = spec->near_log_scale_sd;
    lrt::synth::generate(s, count, positions, shape, sh);
    return LRT_OK;
}

lrt_status lrt_pack_splats(const float* fields, uint32_t count, uint32_t rest_per_colour, float* positions,
                           uint32_t* shape, uint32_t* sh) {
    if (fields == nullptr || positions == nullptr || shape == nullptr) return LRT_INVALID_ARGUMENT;
    if (rest_per_colour != 0 && rest_per_colour != 3 && rest_per_colour != 8 && rest_per_colour != 15)
        return LRT_INVALID_ARGUMENT;
    if (rest_per_colour > 0 && sh == nullptr) return LRT_INVALID_ARGUMENT;
    const uint32_t words = rest_per_colour == 0 ? 1u : (rest_per_colour * 3 + 1) / 2;
    const uint32_t stride = 14 + rest_per_colour * 3;
    for (uint32_t i = 0; i < count; ++i) {
        const float* f = fields + static_cast<uint64_t>(i) * stride;
        lrt::synth::SplatFields s;
        for (int k = 0; k < 3; ++k) s.position[k] = f[k];
        s.opacity = f[3];
        for (int k = 0; k < 3; ++k) s.logScale[k] = f[4 + k];
        for (int k = 0; k < 4; ++k) s.rotation[k] = f[7 + k];
        for (int k = 0; k < 3; ++k) s.base[k] = f[11 + k];
        for (uint32_t k = 0; k < rest_per_colour * 3; ++k) s.rest[k] = f[14 + k];
        lrt::synth::packSplat(s, rest_per_colour, words, positions + uint64_t{i} * 4, shape + uint64_t{i} * 4,
                              sh != nullptr ? sh + uint64_t{i} * words : nullptr);
    }
    return LRT_OK;
}

}   // extern "C"
