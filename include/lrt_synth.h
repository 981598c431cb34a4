/*
 * lrt_synth.h -- host-side splat packing and seeded synthetic scenes.
 *
 * No counterpart in the reference's rasteriser interface: the reference packs
 * splats on the GPU while loading files (scene::CloudLoader::upload,
 * modules/scene/src/GpuClouds.cpp:199-227, over shaders/lrt/scene/splat_decode.slang)
 * and builds test clouds with tests/render/SplatFixtures.h. These calls produce
 * the same packed layout (packing.slang:5-11) in host memory, for benchmarks and
 * tests, with no device needed.
 */
#ifndef LRT_SYNTH_H
#define LRT_SYNTH_H

#include <stdint.h>

#include "lrt_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* kind 0: means uniform in [-1,1]^3 + offset; 1: clusters (centres uniform in
 * [-box_half, box_half]^3, members N(centre, diag(U[sd_lo, sd_hi])^2)); 2: kind 0
 * with near_fraction of the splats pulled toward the origin and drawn with
 * log-scales N(near_log_scale_mean, near_log_scale_sd). Log-scales
 * N(log_scale_mean, log_scale_sd); rotations normalised 4-D normals; opacity
 * sigmoid(N(opacity_logit_mean, opacity_logit_sd)); base rgb U[0,1]; band-l
 * harmonics N(0, sh_sd / l). Every splat draws from its own seeded stream. */
typedef struct {
    uint32_t kind;
    uint32_t degree;
    uint64_t seed;
    float offset[3];
    float log_scale_mean, log_scale_sd;
    float opacity_logit_mean, opacity_logit_sd;
    float sh_sd;
    uint32_t clusters;
    float box_half, cluster_sd_lo, cluster_sd_hi;
    float near_fraction, near_log_scale_mean, near_log_scale_sd;
} lrt_scene_spec;

/* Host arrays: positions count*4 floats, shape count*4 u32, sh count*sh_words
 * u32 (sh_words = (rest*3+1)/2 for rest = 3, 8, 15 at degree 1..3; NULL at
 * degree 0). */
lrt_status lrt_synth_scene(const lrt_scene_spec* spec, uint32_t count, float* positions, uint32_t* shape,
                           uint32_t* sh);

/* Packs count splats given as rows of 14 + 3*rest_per_colour floats: x y z,
 * linear opacity, 3 log-scales, rotation x y z w, base r g b, rest rgb per basis. */
lrt_status lrt_pack_splats(const float* fields, uint32_t count, uint32_t rest_per_colour, float* positions,
                           uint32_t* shape, uint32_t* sh);

#ifdef __cplusplus
}
#endif
#endif
