/*
 * lrt_oracle -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
 *
 * A plain-C, CPU restatement of lucabRTrender's Gaussian-splat tile rasteriser
 * (lrt::render::TileRasterizer) and of its brute-force ground truth
 * (lrt::render::ReferenceRenderer), op for op in FP32, plus an FP64 twin of the
 * brute force. Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg may load it, and only as the checker; the HIP library never links it.
 *
 * Every function names the reference file:line it follows. Where the reference
 * calls a Slang intrinsic whose bits are not pinned by any reference test
 * (exp, log), the oracle defines that function itself (or_expf/or_logf: Cody-Waite
 * range reduction + Taylor polynomial) and the HIP library uses the same
 * published method, so "bit-exact" means bit-exact with this restatement.
 *
 * Parity pins (what checks this oracle beyond retyping it), see tests/:
 *   or_expf/or_logf      vs libm in double, <= 2 ulp over sweeps (test_oracle_math)
 *   half conversions     vs numpy float16 on all 65536 halves / random floats
 *   quaternion packing   round trips + known answers of tests/scene/test_loading.cpp:91-165
 *   record decode        known answers of tests/scene/test_loading.cpp:91-165
 *   projection/camera    closed-form single splat (tests/render/test_splat_render.cpp:123-150),
 *                        orthographic EWA exactness, FP64 twin agreement
 *   scan                 closed form on counts_generate inputs (shaders/lrt/test/counts_generate.slang)
 *   radix sort           brute-force stable sort + sort_generate's 5 patterns
 *                        (shaders/lrt/test/sort_generate.slang, tests/gpu/test_algo.cpp:347-379)
 *   emit/ranges          invariants: each tile's range holds exactly the pairs whose
 *                        pixel-centre box the footprint reaches, depth-ordered
 *   blend                vs the brute force (p99 <= 2, max <= 16 code values,
 *                        tests/render/test_splat_render.cpp:70-121), depth order and
 *                        handedness (:152-200), AA energy (:202-229)
 */
#ifndef LRT_ORACLE_H
#define LRT_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- math the reference takes from Slang intrinsics ------------------- */
float or_expf(float x);
float or_logf(float x);
float or_half_to_float(uint16_t h);
uint16_t or_float_to_half(float f);

/* ---- packing (shaders/lrt/common/packing.slang) ------------------------ */
uint32_t or_encode_quaternion(const float q_xyzw[4]);
void or_decode_quaternion(uint32_t word, float q_xyzw[4]);
uint32_t or_pack_halves(float lo, float hi);

/* A raw record set and how its fields are encoded (modules/io/include/lrt/io/
 * RawSplats.h:15-65): field offsets in floats, rest coefficients colour-outer
 * (rrr.. ggg.. bbb.., a 3DGS PLY) or rgb per basis, and the decode modes,
 * numbered as RawSplats.h:42-58 numbers them:
 *   opacity_mode   0 logit, 1 linear, 2 byte (v / 255)
 *   scale_mode     0 log, 1 linear (log(max(v, 1e-30))), 2 SPZ byte (v / 16 - 10)
 *   colour_mode    0 SH DC (0.5 + C0 v), 1 linear, 2 byte (v / 255),
 *                  3 SPZ byte (0.5 + C0 ((v / 255 - 0.5) / 0.15))
 *   rotation_mode  0 float (w, x, y, z), 1 byte ((v - 128) / 128 each),
 *                  2 SPZ v2 (x y z bytes v / 127.5 - 1, w from unit length),
 *                  3 SPZ v3 smallest three (a u32: low 16 bits at rot_x, high 16 at rot_y)
 *   rest_mode      0 float, 1 byte ((v - 128) / 128)
 * Positions are multiplied by 2^-fractional_bits; flip_yz: the record is
 * right-up-back (SPZ) and y and z change sign in positions and rotations, and
 * in the rest coefficients odd in y xor z (splat_encoding.slang:35-112).
 * All zero is a 3DGS PLY (tests/render/SplatFixtures.h CloudBuilder). */
typedef struct {
    uint32_t floats_per_record;
    uint32_t x, y, z, opacity, scale0, scale1, scale2, rot_w, rot_x, rot_y, rot_z;
    uint32_t dc0, dc1, dc2, rest_base, file_per_colour, rest_colour_outer;
    uint32_t opacity_mode, scale_mode, colour_mode, rotation_mode, rest_mode;
    uint32_t fractional_bits, flip_yz;
} or_encoding;

/* splat_validate + scan + splat_decode for one slice (shaders/lrt/scene/
 * splat_validate.slang:16-29, splat_decode.slang:17-87, splat_encoding.slang:
 * 28-96, GpuClouds.cpp:151-227). Writes compacted packed splats; returns the
 * number kept. */
uint32_t or_decode_records(const float* records, uint32_t count, const or_encoding* e,
                           uint32_t keep_per_colour, float* positions, uint32_t* shape,
                           uint32_t* sh, uint32_t sh_words);

/* ---- camera and frame parameters (host, FP64) ------------------------- */
typedef struct {
    double camera_to_world[16];          /* row-major 4x4 */
    int32_t orthographic;
    double focal, haperture, vaperture, near_z, far_z;
    double window_translate[2], window_scale[2], window_roll;
} or_camera;

typedef struct {
    double world_to_view[16];
    double focal_x, focal_y, centre_x, centre_y, near_z, far_z;
    int32_t orthographic;
    double eye_world[3];
} or_projection;

typedef struct {
    uint32_t active, shape, mode, invert;
    float centre[3], size[3], tint[3];
    float saturation, brightness, opacity, min_opacity, max_scale;
} or_edit;

/* The kernel constant block (shaders/lrt/splat/frame.slang:33-81), trimmed to
 * the fields the splat-only path reads, plus the two filter parameters this
 * build adds (dilation: the reference's fixed 0.3; use_filter3d: Mip-Splatting
 * 3D filter variance per splat). */
typedef struct {
    uint32_t width, height, tiles_x, tiles_y;
    uint32_t tile_shift;   /* log2 of the tile's side: 3, 4 (the reference's 16 pixels) or 5 */
    float focal_x, focal_y, centre_x, centre_y, near_z, far_z;
    /* The Jacobian's clamp, 1.3x the frustum in units of x/z and y/z: the
     * expressions the projection evaluates (frame.slang's splatProject
     * clamp), kept here so the frame block matches the library's byte for byte. */
    float lim_lo_x, lim_hi_x, lim_lo_y, lim_hi_y;
    uint32_t orthographic, antialias;
    float m[12];
    float eye[3];
    uint32_t count, base, rest_per_colour, sh_words, sh_limit, tile_bits;
    uint32_t depth_mode;
    float depth_threshold;
    float bg[4];
    uint32_t linearise;
    float dilation;
    uint32_t use_filter3d;
    or_edit edit;
} or_frame;

typedef struct {
    uint32_t width, height;
    int32_t antialias;
    uint32_t max_sh_degree;
    float background[4];
    int32_t depth_mode;          /* 0 mean, 1 threshold */
    float depth_threshold;
    int32_t linearise;
    float dilation;              /* 0.3 reproduces the reference */
    int32_t filter3d;            /* 0 reproduces the reference */
    uint32_t tile_size;          /* 8, 16 or 32 pixels a tile's side; 0 or 16 reproduces the reference */
} or_settings;

void or_look_at(const double eye[3], const double target[3], const double up[3], double out[16]);
void or_mat_mul(const double a[16], const double b[16], double out[16]);
void or_inverse_affine(const double m[16], double out[16]);
void or_translation(const double t[3], double out[16]);
void or_scaling(const double s[3], double out[16]);
void or_rotation_y(double degrees, double out[16]);
void or_rotation_z(double degrees, double out[16]);
void or_projection_for(const or_camera* camera, uint32_t width, uint32_t height, or_projection* out);
uint32_t or_bits_for(uint32_t values);
/* setFrame + setObject (modules/render/src/FrameParams.h:32-69, TileRasterizer.cpp:297-314). */
void or_frame_for(const or_projection* p, const or_settings* s, const double object_to_world[16],
                  const or_edit* edit, uint32_t count, uint32_t base, uint32_t rest_per_colour,
                  uint32_t sh_words, or_frame* out);

/* ---- the rasteriser's stages ----------------------------------------- */
typedef struct { float screen[4]; float conic[4]; float colour[4]; } or_proj_record;

/* splatProject (shaders/lrt/splat/splat_project.slang:52-221), frame.count splats
 * from index 0 of the arrays, written to slots frame.base + i. */
void or_project(const or_frame* f, const float* positions, const uint32_t* shape, const uint32_t* sh,
                const float* filter3d, or_proj_record* proj, uint32_t* rects, uint32_t* touched,
                uint32_t* visible, uint32_t* depth_keys, int threads);

uint64_t or_scan_exclusive(const uint32_t* in, uint32_t* out, uint32_t n);
/* Stable LSD radix sort of (key, value) on the low key_bits bits, 8 bits a pass
 * (modules/gpu/src/algo/RadixSort.cpp:80-158 semantics). 64-bit keys. */
void or_radix_sort(uint64_t* keys, uint32_t* values, uint32_t n, uint32_t key_bits);

/* The whole frame, as TileRasterizer::render orders it (TileRasterizer.cpp:188-484)
 * for a set of instances already turned into frames (frames[k].base set).
 * Outputs (each may be NULL): sorted pair tiles and splats (P of each), ranges
 * (2 per tile), image (4 floats per pixel) and depth. Per-splat arrays are sized
 * total = sum of frames[k].count. Returns P. */
typedef struct {
    const float* positions;
    const uint32_t* shape;
    const uint32_t* sh;
    const float* filter3d;
} or_cloud;

typedef struct {
    uint32_t splats, visible, pairs;
} or_stats;

uint32_t or_raster(const or_frame* frames, const or_cloud* clouds, uint32_t instances,
                   or_proj_record* proj, uint32_t* touched, uint32_t* depth_keys,
                   uint32_t* pair_tiles, uint32_t* pair_splats, uint32_t pair_capacity,
                   uint32_t* ranges, float* colour, float* depth, uint32_t* last_count,
                   float* final_t, or_stats* stats, int threads);

/* ---- points as opaque discs ------------------------------------------- */
/* render::PointStyle (modules/render/include/lrt/render/Points.h:21-33), the
 * fields the disc route reads (point_frame.slang:7-29). */
typedef struct {
    uint32_t size_mode;          /* 0 world units, 1 pixels */
    float size;                  /* diameter */
    int32_t constant_colour;     /* 1: colour below; 0: the cloud's own */
    float colour[3];
} or_point_style;

/* A point instance (render::PointInstance, Points.h:35-39) whose frame is set
 * apart: positions float4 xyz 1, colours 2 u32 a point (f16 r|g, b|a). */
typedef struct {
    const float* positions;
    const uint32_t* colours;
    or_point_style style;
} or_points;

/* pointsProject (shaders/lrt/splat/points_project.slang:33-86), frame.count
 * points from index 0, written to slots frame.base + i. */
void or_points_project(const or_frame* f, const or_point_style* style, const float* positions,
                       const uint32_t* colours, or_proj_record* proj, uint32_t* rects, uint32_t* touched,
                       uint32_t* visible, uint32_t* depth_keys, int threads);

/* TileRasterizer::render with point instances and an under-layer
 * (TileRasterizer.cpp:188-484): the splat instances' slots first, then the
 * points'; the composite blend. under_colour (4 floats a pixel) and
 * under_depth may be NULL (no under-layer). */
uint32_t or_raster_composite(const or_frame* frames, const or_cloud* clouds, uint32_t instances,
                             const or_frame* point_frames, const or_points* points, uint32_t point_instances,
                             const float* under_colour, const float* under_depth, or_proj_record* proj,
                             uint32_t* touched, uint32_t* depth_keys, uint32_t* pair_tiles, uint32_t* pair_splats,
                             uint32_t pair_capacity, uint32_t* ranges, float* colour, float* depth,
                             or_stats* stats, int threads);

/* splatBlendComposite (shaders/lrt/splat/splat_blend.slang:23-108,116-120). */
void or_blend_composite(const or_frame* f, const uint32_t* ranges, const uint32_t* pair_splats,
                        const or_proj_record* proj, const float* under_colour, const float* under_depth,
                        float* colour, float* depth, int threads);

/* splatBlend (shaders/lrt/splat/splat_blend.slang:23-114), splat-only entry. */
void or_blend(const or_frame* f, const uint32_t* ranges, const uint32_t* pair_splats,
              const or_proj_record* proj, float* colour, float* depth, uint32_t* last_count,
              float* final_t, int threads);

/* ---- the ground truth ------------------------------------------------- */
/* referenceProject + referenceBlend (shaders/lrt/reference/reference_project.slang:31-105,
 * reference_blend.slang:24-119) in FP32, without the 1024-hit cap: every splat
 * in (view z, index) order is walked per pixel. px/py select a pixel subset
 * (NULL: every pixel, colour/depth full images); otherwise colour/depth hold
 * one entry per listed pixel. keyed = 1 orders by the rasteriser's 24-bit
 * depth key instead of the exact view z (ties by index either way). */
void or_reference_render_f32(const or_frame* frames, const or_cloud* clouds, uint32_t instances,
                             const uint32_t* px, const uint32_t* py, uint32_t npix, float* colour,
                             float* depth, int keyed, int threads);
/* The brute force with point instances as discs (reference_points_project.slang:
 * 20-42; reference_blend.slang:50-51,80-115): alpha 1, and the nearest disc
 * reaching a pixel ends its walk. Whole images, (view z, index) order. */
void or_reference_render_points_f32(const or_frame* frames, const or_cloud* clouds, uint32_t instances,
                                    const or_frame* point_frames, const or_points* points,
                                    uint32_t point_instances, float* colour, float* depth, int threads);
/* The same in FP64 throughout (the error oracle). */
void or_reference_render_f64(const or_frame* frames, const or_cloud* clouds, uint32_t instances,
                             const uint32_t* px, const uint32_t* py, uint32_t npix, double* colour,
                             double* depth, int keyed, int threads);

/* A cloud of FP32 parameters (the differentiable frame's input): means N x 3,
 * log scales N x 3, quaternions N x 4 (x, y, z, w; normalised on use), opacity
 * logits N, harmonics N x (1 + rest_per_colour) x 3 (dc first), filter3d N or
 * NULL. */
typedef struct {
    const float *means, *log_scales, *quats, *logits, *sh, *filter3d;
    uint32_t count, rest_per_colour;
} or_diff_cloud;
/* The brute force of such a cloud in FP64 (frame.base must be 0): scales
 * exp(log s), opacity sigmoid(logit), base colour 0.5 + C0 dc, then
 * reference_project.slang:31-105 and the walk of reference_blend.slang:24-119. */
void or_diff_render_f64(const or_frame* frame, const or_diff_cloud* cloud, const uint32_t* px, const uint32_t* py,
                        uint32_t npix, double* colour, double* depth, int keyed, int threads);

/* ---- levels of detail (modules/lod/src/Lod.cpp, shaders/lrt/lod) ------- */
/* The 30-bit Morton code of a point at unit coordinates (morton.slang:27-30). */
uint32_t or_morton30(float ux, float uy, float uz);
/* lodMorton: keys of positions (float4) in the cube [lo, lo + extent]^3, values = index. */
void or_lod_morton(const float* positions, uint32_t n, const float lo[3], float extent, uint64_t* keys,
                   uint32_t* values);
/* lodBoundaries + PrefixSum + lodGroups at `level` over Morton-sorted keys:
 * group a splat, starts/cells a group (any may be NULL). Returns the groups. */
uint32_t or_lod_groups(const uint64_t* keys, uint32_t n, uint32_t level, uint32_t* group, uint32_t* starts,
                       uint32_t* cells);
/* lodLeafMoments / lodMergeMoments: 13 + 3 keep floats a group. */
void or_lod_leaf_moments(const float* positions, const uint32_t* shape, const uint32_t* sh, uint32_t n,
                         const uint32_t* starts, uint32_t groups, uint32_t keep, uint32_t sh_words, float* moments);
void or_lod_merge_moments(const uint32_t* starts, uint32_t groups, const uint32_t* fine_group, uint32_t fine_groups,
                          const float* fine_moments, uint32_t stride, float* moments);
/* lodFinalize: one packed Gaussian a group. */
void or_lod_finalize(const float* moments, uint32_t groups, uint32_t keep, uint32_t sh_words, float* positions,
                     uint32_t* shape, uint32_t* sh);
/* The cut's view of a cloud (LodParams, lod_common.slang:40-62, and
 * Lod.cpp:410-433): the bounds, the eye in the cloud's space, pixels a unit. */
typedef struct {
    float lo[3];
    float extent;
    float eye[3];
    float pixels_per_unit;
    float threshold;
    uint32_t orthographic;
} or_cut_params;
float or_lod_projected_edge(const or_cut_params* p, uint32_t cell, uint32_t level);
void or_lod_cut_groups(const or_cut_params* p, const uint32_t* cells, uint32_t count, uint32_t level, int coarsest,
                       uint32_t* selected);
void or_lod_cut_splats(const or_cut_params* p, const uint32_t* groups, uint32_t n, const uint32_t* finest_cells,
                       uint32_t finest_level, uint32_t* selected);

/* compareImages (modules/render/src/ReferenceRenderer.cpp:248-297 with
 * shaders/lrt/reference/image_compare.slang:27-51): 8-bit sRGB code-value
 * histogram of the worst channel; p99, max, pixels over 2. */
void or_compare_images(const float* a, const float* b, uint32_t width, uint32_t height,
                       uint32_t* p99, uint32_t* max, uint64_t* over2);

/* The reference's sRGB curve (shaders/lrt/common/color.slang:7-10). */
float or_srgb_to_linear(float c);

#ifdef __cplusplus
}
#endif
#endif
