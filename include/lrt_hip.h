/*
 * lrt_hip.h -- the C ABI of the MI355X Gaussian-splat tile rasteriser.
 *
 * The reference has no C ABI: its rasteriser is the C++ class
 * lrt::render::TileRasterizer over slang-rhi. Each entry point below names the
 * reference interface it replaces (path:line in lucabRTrender). All pointers
 * called "device" are HIP device pointers the caller owns (hipMalloc, or a
 * torch tensor's data_ptr()); "host" pointers are ordinary memory. No call
 * takes a torch type. Every call that can fail returns an lrt_status; the
 * message of the last failure on a context is lrt_last_error(ctx).
 *
 * Threading: one context is one caller and one HIP stream, as the reference's
 * device has one caller (CLAUDE.md:110-112). Calls on one context must not
 * overlap; different contexts are independent.
 */
#ifndef LRT_HIP_H
#define LRT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LRT_ABI_VERSION 1

/* lrt::ErrorCode (modules/core/include/lrt/core/Error.h:14-26), shifted so
 * that 0 is success. */
typedef enum {
    LRT_OK = 0,
    LRT_UNKNOWN = 1,
    LRT_NOT_FOUND = 2,
    LRT_INVALID_ARGUMENT = 3,
    LRT_UNSUPPORTED = 4,
    LRT_IO_FAILURE = 5,
    LRT_DEVICE_FAILURE = 6,
    LRT_SHADER_FAILURE = 7,
    LRT_PLUGIN_FAILURE = 8,
    LRT_OUT_OF_MEMORY = 9,
    LRT_CANCELLED = 10,
    LRT_INTERNAL_ERROR = 11
} lrt_status;

typedef struct lrt_context lrt_context;

/* ---- host math: cameras (no device needed) ------------------------------ */

/* lrt::render::Lens (modules/render/include/lrt/render/Camera.h:23-44), the
 * fields the rasteriser reads. lrt_lens_default fills the reference's defaults. */
typedef struct {
    int32_t orthographic;
    double focal;            /* mm */
    double haperture;        /* mm */
    double vaperture;        /* mm */
    double near_z;
    double far_z;
    double window_translate[2];
    double window_scale[2];
    double window_roll;      /* degrees */
} lrt_lens;

/* lrt::render::Camera (Camera.h:46-53): object-to-world, row-major 4x4. */
typedef struct {
    double camera_to_world[16];
    lrt_lens lens;
} lrt_camera;

/* lrt::render::Projection (Camera.h:56-80), the fields the rasteriser reads. */
typedef struct {
    double world_to_view[16];
    double focal_x, focal_y, centre_x, centre_y, near_z, far_z;
    int32_t orthographic;
    double eye_world[3];
} lrt_projection;

void lrt_lens_default(lrt_lens* out);
/* Camera::lookingAt (modules/render/src/Camera.cpp:9-13, aofx Transform.h:376-400). */
void lrt_camera_looking_at(const double eye[3], const double target[3], const double up[3], lrt_camera* out);
/* projectionFor (modules/render/src/Camera.cpp:20-48). */
lrt_status lrt_projection_for(const lrt_camera* camera, uint32_t width, uint32_t height, lrt_projection* out);

/* ---- the scene ------------------------------------------------------------ */

/* scene::GpuSplats (modules/scene/include/lrt/scene/GpuClouds.h:41-76), packed as
 * shaders/lrt/common/packing.slang:5-11 lays it out. Device pointers.
 *   positions  float4 a splat: x, y, z, linear opacity
 *   shape      4 u32 a splat: smallest-three rotation; f16 ln sx | ln sy;
 *              f16 ln sz | base r; f16 base g | base b
 *   sh         sh_words u32 a splat: f16 rest coefficients, rgb per basis
 *              (may be NULL when rest_per_colour is 0)
 *   filter3d   float a splat: Mip-Splatting 3D filter std (NULL: none); read
 *              only when lrt_render_settings.filter3d is set. Not in the reference. */
typedef struct {
    const float* positions;
    const uint32_t* shape;
    const uint32_t* sh;
    const float* filter3d;
    uint32_t count;
    uint32_t rest_per_colour;   /* 0, 3, 8, 15 */
    uint32_t sh_words;          /* (rest_per_colour * 3 + 1) / 2, 1 at degree 0 */
} lrt_splats;

/* render::SplatEdit (modules/render/include/lrt/render/TileRasterizer.h:29-50). */
typedef struct {
    int32_t active;
    uint32_t shape;      /* 0 box, 1 sphere */
    uint32_t mode;       /* 0 keep inside, 1 remove inside, 2 grade inside */
    int32_t invert;
    float centre[3];
    float size[3];
    float tint[3];
    float saturation, brightness, opacity, min_opacity, max_scale;
} lrt_splat_edit;

/* render::SplatInstance (TileRasterizer.h:52-67), without relighting. */
typedef struct {
    const lrt_splats* splats;
    double object_to_world[16];
    lrt_splat_edit edit;
} lrt_splat_instance;

/* render::RenderSettings (TileRasterizer.h:105-118) plus two filter settings the
 * reference fixes: dilation (its +0.3 px^2) and filter3d (off). */
typedef struct {
    uint32_t width, height;
    int32_t antialias;
    uint32_t max_sh_degree;
    float background[4];          /* premultiplied, linear */
    int32_t depth_mode;           /* 0 mean, 1 threshold */
    float depth_threshold;
    int32_t linearise;
    int32_t time_stages;          /* per-stage HIP event timing into lrt_frame_stats */
    float dilation;               /* 0.3 reproduces the reference */
    int32_t filter3d;             /* 1: read lrt_splats.filter3d */
    uint32_t blend_batch;         /* records each 8x8 quadrant's LDS list holds: 64, 128 or 256 (0, the
                                     default: 256; other values take the next of those up, 256 at most) */
    int32_t tile_order;           /* 0: tile index order; 1 (default): heaviest tiles launched first */
    int32_t persistent_blend;     /* batches: 1 = a persistent blend taking quadrants from a queue */
    int32_t use_graph;            /* batches: 1 = capture the launch sequence once and replay it */
    uint32_t blend_segment;       /* a tile with more pairs is split into segments walked by waves of their
                                     own and composited, a pixel that stops inside a segment walking it
                                     again (0, the default: never). Frames without points, an under-layer
                                     or a backward, with tile_order 1 */
    int32_t keep_walk_state;      /* 1: keep each pixel's last contributor and final transmittance for
                                     LRT_DEBUG_LAST_COUNT / LRT_DEBUG_FINAL_T (a differentiable frame
                                     always keeps them; otherwise they cost the blend an instruction a pair) */
    uint32_t tile_size;           /* pixels a square tile's side: 8, 16 or 32 (0, the default: 16, the
                                     reference's frame.slang:31). Differentiable frames take 16 only */
    uint32_t radix_bits;          /* the sort's widest digit: 8, 9 or 10 bits (0, the default: 9 where that
                                     saves a pass, else 8); the sorted pairs are the same */
} lrt_render_settings;

/* render::RenderTargets (TileRasterizer.h:129-134): device images, bottom row
 * first, width * height each. colour: float4 premultiplied linear; depth:
 * float view z (may be NULL). The caller allocates them at the settings' size. */
typedef struct {
    uint32_t width, height;
    float* colour;
    float* depth;
} lrt_render_targets;

/* render::FrameStats (TileRasterizer.h:120-126). depth_sort_ms is 0: one
 * 64-bit (tile, depth) sort replaces the reference's depth sort + tile sort
 * and is reported as tile_sort_ms. total_ms is the call's wall clock, as the
 * reference's; the stage figures (with time_stages) are GPU time from HIP
 * events, counts_ms including the pair-count readback's wait. key_bits and
 * sort_passes (no counterpart in the reference) describe the frame's sort: its
 * virtual key is the tile id over the depth key (bits(view z) >> 7) less the
 * frame's least drawn one, in bitsFor(tiles) + bitsFor(depth-key range) bits,
 * eight a pass (a batch of views, which reads no counts back before its sort,
 * bounds the depth keys by its views' least near and greatest far plane). */
typedef struct {
    uint32_t splats, visible, pairs;
    double project_ms, depth_sort_ms, counts_ms, emit_ms, tile_sort_ms, blend_ms, total_ms;
    uint32_t key_bits, sort_passes;
} lrt_frame_stats;

void lrt_render_settings_default(lrt_render_settings* out);
void lrt_splat_edit_default(lrt_splat_edit* out);

/* ---- the rasteriser ---------------------------------------------------------- */

/* TileRasterizer::create (TileRasterizer.h:138, TileRasterizer.cpp:59-103).
 * device: HIP device ordinal; stream: a hipStream_t to queue on (NULL: the
 * null stream). Nothing is compiled at run time. */
lrt_status lrt_context_create(int device, void* stream, lrt_context** out);
void lrt_context_destroy(lrt_context* ctx);
lrt_status lrt_context_set_stream(lrt_context* ctx, void* stream);
/* The message of the context's last failure; with NULL, that of the calling
 * thread's last call that takes no context (lrt_context_create,
 * lrt_splats_read, lrt_raw_splats_create, lrt_write_exr), empty when it
 * succeeded. */
const char* lrt_last_error(const lrt_context* ctx);

/* TileRasterizer::render(const Projection&, ...) (TileRasterizer.h:143-155,
 * TileRasterizer.cpp:188-484), splat instances only. Synchronous, as the
 * reference's: returns once the images are written. stats may be NULL. */
lrt_status lrt_raster_render(lrt_context* ctx, const lrt_projection* projection,
                             const lrt_splat_instance* instances, uint32_t instance_count,
                             const lrt_render_settings* settings, lrt_render_targets* targets,
                             lrt_frame_stats* stats);

/* ---- points as opaque discs, and the under-layer ------------------------------------ */

/* scene::GpuPoints (modules/scene/include/lrt/scene/GpuClouds.h:78-85). Device
 * pointers.
 *   positions  float4 a point: x, y, z, 1
 *   colours    2 u32 a point: f16 r | g << 16, b | a << 16 (linear), 8-byte
 *              aligned; may be NULL when the style's constant_colour is set */
typedef struct {
    const float* positions;
    const uint32_t* colours;
    uint32_t count;
} lrt_points;

/* render::PointStyle (modules/render/include/lrt/render/Points.h:21-33), the
 * fields the disc route reads; the eye-dome and surface-splatting fields
 * belong to the reference's raster route (PointRasterizer), which this
 * library does not have. lrt_point_style_default fills the reference's
 * defaults: world units, 0.01, the cloud's own colours, white. */
typedef struct {
    uint32_t size_mode;          /* 0 world units, 1 pixels */
    float size;                  /* diameter */
    int32_t constant_colour;     /* 1: colour for every point */
    float colour[3];
} lrt_point_style;

/* render::PointInstance (Points.h:35-39). */
typedef struct {
    const lrt_points* points;
    double object_to_world[16];
    lrt_point_style style;
} lrt_point_instance;

void lrt_point_style_default(lrt_point_style* out);

/* TileRasterizer::render(projection, instances, settings, targets, points,
 * under) (TileRasterizer.h:143-155, TileRasterizer.cpp:188-484 with :355-377
 * and :464-475): point instances drawn as opaque discs
 * (points_project.slang:33-86) in the same depth order as the splats, their
 * slots after the splats'; `under` (may be NULL; the frame's size, colour and depth both set,
 * device memory as the targets') is what lies under the splats -- each
 * pixel's walk ends at the first pair at or behind its under depth (when that
 * is above 0) and the under colour lies behind. Points or an under-layer
 * select the composite blend (splat_blend.slang:23-120); with neither this is
 * lrt_raster_render. A composite frame keeps no per-pixel walk state
 * (LRT_DEBUG_LAST_COUNT and LRT_DEBUG_FINAL_T are empty after it).
 * Synchronous; stats may be NULL. */
lrt_status lrt_raster_render_composite(lrt_context* ctx, const lrt_projection* projection,
                                       const lrt_splat_instance* instances, uint32_t instance_count,
                                       const lrt_point_instance* points, uint32_t point_count,
                                       const lrt_render_targets* under, const lrt_render_settings* settings,
                                       lrt_render_targets* targets, lrt_frame_stats* stats);

/* ---- levels of detail ----------------------------------------------------------------
 * lrt::lod (modules/lod/include/lrt/lod/Lod.h:40-143; shaders/lrt/lod/ kernels)
 * for a cloud held whole in device memory: the reference's streamed store
 * (chunks, residency, .lrtc) is not carried. The build sorts the splats by the
 * Morton code of their position in the bounding cube and keeps, for each
 * stored octree level (coarsest .. finest, consecutive), one merged Gaussian a
 * non-empty cell from the splats' additive moments. The cut draws a merged
 * Gaussian when its cell projects to at most `threshold` pixels and its
 * parent's does not, and a splat when its finest-level cell does not:
 * threshold 0 draws every splat and nothing merged. */
typedef struct lrt_lod_cloud lrt_lod_cloud;

/* LodBuildSettings (Lod.h:78-87) without the streaming chunk size. */
typedef struct {
    float max_group_fraction;    /* the finest stored level's cells hold >= 1 / this splats on average */
    uint32_t coarsest_level;     /* 1 .. 10 */
} lrt_lod_build_settings;

void lrt_lod_build_settings_default(lrt_lod_build_settings* out);

/* LodBuilder::build (Lod.h:89-99, Lod.cpp:81-283) on the context's stream.
 * Synchronous. The cloud (device pointers, as for lrt_raster_render) is read
 * only; the levels own their device memory until lrt_lod_destroy. settings
 * may be NULL (the defaults). An empty cloud is LRT_INVALID_ARGUMENT. */
lrt_status lrt_lod_build(lrt_context* ctx, const lrt_splats* cloud, const lrt_lod_build_settings* settings,
                         lrt_lod_cloud** out);
void lrt_lod_destroy(lrt_lod_cloud* lod);

/* What a built cloud holds (LodCloud, Lod.h:51-76): per stored level, coarsest
 * first, its octree level, its first merged Gaussian and its group count. */
typedef struct {
    uint32_t count, rest_per_colour, sh_words;
    uint32_t levels, merged;
    uint32_t level[10], first[10], groups[10];
    float bounds_lo[3];
    float extent;                /* the bounds' largest edge x 1.0001: cells are cubes */
} lrt_lod_info;
lrt_status lrt_lod_get_info(const lrt_lod_cloud* lod, lrt_lod_info* out);
/* The cloud's splats in Morton order and (groups, may be NULL) each one's
 * finest-level group; level `level_index`'s merged Gaussians and (cells, may
 * be NULL) their cell codes. Device pointers the cloud owns. */
lrt_status lrt_lod_splats(const lrt_lod_cloud* lod, lrt_splats* out, const uint32_t** groups);
lrt_status lrt_lod_gaussians(const lrt_lod_cloud* lod, uint32_t level_index, lrt_splats* out,
                             const uint32_t** cells);

/* LodInstance (Lod.h:101-106): has_threshold 1 uses `threshold` over the call's. */
typedef struct {
    const lrt_lod_cloud* cloud;
    double object_to_world[16];
    lrt_splat_edit edit;
    int32_t has_threshold;
    float threshold;
} lrt_lod_instance;

/* CutStats (Lod.h:108-117) without the streaming needs. */
typedef struct {
    uint32_t splats, merged, available;
} lrt_cut_stats;

/* CutSelector::select (Lod.h:131-134, Lod.cpp:350-547): each instance's cut
 * for this view as splat instances for lrt_raster_render, in order, an
 * instance that draws nothing left out. out has room for `count` entries and
 * *out_count gets how many were written; stats (may be NULL) gets one entry an
 * instance. The instances' clouds are device memory the context owns, valid
 * until its next lrt_lod_select; their gathers are queued on the context's
 * stream after one wait for the drawn counts. */
lrt_status lrt_lod_select(lrt_context* ctx, const lrt_projection* projection, const lrt_lod_instance* instances,
                          uint32_t count, float threshold, lrt_splat_instance* out, uint32_t* out_count,
                          lrt_cut_stats* stats);

/* ---- the brute force in FP64 -------------------------------------------------------
 * ReferenceRenderer (modules/render/src/ReferenceRenderer.cpp:54-160;
 * reference_project.slang:31-105, reference_blend.slang:24-119) on the GPU in
 * double, without its per-pixel hit cap: every splat of every instance in
 * (view z, index) order, or with keyed in the rasteriser's (24-bit depth key,
 * slot) order, walked per pixel. colour: 4 doubles a pixel, depth: 1 (may be
 * NULL), device memory of the settings' size, bottom row first. Synchronous.
 * It clobbers nothing the rasteriser's debug buffers hold. */
lrt_status lrt_reference_render(lrt_context* ctx, const lrt_projection* projection,
                                const lrt_splat_instance* instances, uint32_t instance_count,
                                const lrt_render_settings* settings, double* colour, double* depth,
                                int32_t keyed);
/* The same brute force on a subset of the frame's pixels (the reference's
 * renderer draws whole frames; a subset is what a full-size check can afford):
 * pixels (device, pixel_count u32, each y * width + x, bottom row first);
 * colour receives 4 doubles and depth (may be NULL) 1 for each listed pixel,
 * in list order. */
lrt_status lrt_reference_render_pixels(lrt_context* ctx, const lrt_projection* projection,
                                       const lrt_splat_instance* instances, uint32_t instance_count,
                                       const lrt_render_settings* settings, const uint32_t* pixels,
                                       uint32_t pixel_count, double* colour, double* depth, int32_t keyed);

/* ---- batches of views -------------------------------------------------------------
 * No single counterpart in the reference, which draws one camera a render()
 * call (TileRasterizer.h:143-155): the boundary's lrt_raster_render_views. */

/* One camera of a batch: its projection and its images (targets.width/height
 * are the view's frame size). */
typedef struct {
    lrt_projection projection;
    lrt_render_targets targets;
} lrt_view;

/* Every view of the same splat instances in one launch sequence with one wait:
 * each view's images are those lrt_raster_render would draw for it alone, bit
 * for bit. settings.width/height are ignored. stats: the batch's totals
 * (splats counts every view's slots). With settings.use_graph the sequence is
 * captured as a hipGraph once per batch shape and replayed. */
lrt_status lrt_raster_render_views(lrt_context* ctx, const lrt_view* views, uint32_t view_count,
                                   const lrt_splat_instance* instances, uint32_t instance_count,
                                   const lrt_render_settings* settings, lrt_frame_stats* stats);

/* The autotuner: times the batch under each launch choice this build offers
 * (tile size 8/16/32, list length 64/128/256, grid or persistent blend, graph
 * replay or not), then the sort's digit width (8/9/10 bits) under the fastest,
 * and writes the winner into *settings (its other fields are kept). ms (host,
 * may be NULL) receives the winner's batch time. */
lrt_status lrt_autotune_views(lrt_context* ctx, const lrt_view* views, uint32_t view_count,
                              const lrt_splat_instance* instances, uint32_t instance_count,
                              lrt_render_settings* settings, float* ms);

/* ---- the differentiable frame ---------------------------------------------------
 * No counterpart in the reference, which has no backward: the entries that the
 * boundary's lrt_raster_backward stands for. One cloud of FP32 parameters in
 * device memory, drawn by the same stages as lrt_raster_render; the context
 * keeps what the backward needs until its next render. The backward reads the
 * parameter buffers (and filter3d) again: they must stay allocated and
 * unchanged from the differentiable render to its backward. */
typedef struct {
    const float* means;           /* N x 3, the cloud's own space */
    const float* log_scales;      /* N x 3: scale = exp(log scale) */
    const float* quats;           /* N x 4: x, y, z, w (w real), normalised on use */
    const float* opacity_logits;  /* N: opacity = sigmoid(logit) */
    const float* sh;              /* N x (1 + rest_per_colour) x 3: dc, then the rest coefficients,
                                     rgb per basis; colour = max(0.5 + C0 dc + rest(dir), 0) */
    const float* filter3d;        /* N, or NULL; read when settings.filter3d */
    uint32_t count;
    uint32_t rest_per_colour;     /* 0, 3, 8, 15 */
} lrt_diff_splats;

/* Device gradient buffers of the same shapes, overwritten by the backward. */
typedef struct {
    float* means;
    float* log_scales;
    float* quats;
    float* opacity_logits;
    float* sh;
} lrt_diff_grads;

/* A frame of one FP32-parameter cloud; as lrt_raster_render otherwise. */
lrt_status lrt_raster_render_diff(lrt_context* ctx, const lrt_projection* projection,
                                  const lrt_diff_splats* splats, const double object_to_world[16],
                                  const lrt_render_settings* settings, lrt_render_targets* targets,
                                  lrt_frame_stats* stats);
/* Gradients of the last lrt_raster_render_diff's parameters from dL/d(colour)
 * (device, one float4 a pixel, bottom row first). The depth image receives no
 * gradient. millis (host, 2 floats, may be NULL): GPU ms of the blend's and
 * the preprocess's backward. Synchronous. */
lrt_status lrt_raster_backward(lrt_context* ctx, const float* grad_colour, const lrt_diff_grads* grads,
                               float* millis);

/* Many views of one FP32-parameter cloud in one launch sequence with one wait
 * (no counterpart in the reference, which draws one camera a render() call and
 * has no backward): each view's images are those lrt_raster_render_diff would
 * draw for it alone, bit for bit, and the context keeps what
 * lrt_raster_backward_views needs until its next render. settings.width/height
 * are ignored (each view's targets give its size); settings.use_graph captures
 * the sequence as a hipGraph once per batch shape and replays it. */
lrt_status lrt_raster_render_diff_views(lrt_context* ctx, const lrt_view* views, uint32_t view_count,
                                        const lrt_diff_splats* splats, const double object_to_world[16],
                                        const lrt_render_settings* settings, lrt_frame_stats* stats);
/* Gradients of the last lrt_raster_render_diff_views' parameters, summed over
 * its views in view order (no atomics): grad_colours is a host array of
 * view_count device pointers, view v's dL/d(colour) (one float4 a pixel of
 * that view, bottom row first). view_count must be the batch's. millis as
 * lrt_raster_backward's. Synchronous. */
lrt_status lrt_raster_backward_views(lrt_context* ctx, const float* const* grad_colours, uint32_t view_count,
                                     const lrt_diff_grads* grads, float* millis);

/* What the last render left on the device, copied to host, for bit-exact
 * checks. *needed (may be NULL) receives the byte size; bytes may be 0 to ask. */
typedef enum {
    LRT_DEBUG_COUNTS = 0,        /* u32[4]: splats, visible, pairs, tiles */
    LRT_DEBUG_TOUCHED = 1,       /* u32 a slot: tiles touched (0: culled) */
    LRT_DEBUG_OFFSETS = 2,       /* u32 a slot: exclusive scan of TOUCHED */
    LRT_DEBUG_PROJ = 3,          /* 12 floats a slot: screen, conic, colour */
    LRT_DEBUG_KEYS = 4,          /* u64 a pair, sorted: tile << 32 | bits(view z) */
    LRT_DEBUG_VALUES = 5,        /* u32 a pair, sorted: slot */
    LRT_DEBUG_RANGES = 6,        /* u32[2] a tile: [start, end) */
    LRT_DEBUG_LAST_COUNT = 7,    /* u32 a pixel: pairs walked to the last contributor (kept frames only) */
    LRT_DEBUG_FINAL_T = 8,       /* float a pixel: final transmittance (kept frames only) */
    LRT_DEBUG_RECTS = 9,         /* u32[2] a slot: x0 | y0 << 16, x1 | y1 << 16 */
    LRT_DEBUG_KERNEL_MS = 10,    /* float[8], last render with time_stages: GPU ms of project,
                                    scan, duplicate, sort histogram, sort passes, ranges (with
                                    the heaviest-first tile ordering), blend, and their sum (HIP
                                    events on the context's stream) */
    LRT_DEBUG_TILE_ORDER = 11,   /* u32 a tile: the order the blend took the tiles in (empty
                                    when the last render used index order) */
    LRT_DEBUG_WALK_STEPS = 12    /* u32[1]: the last lrt_raster_render's blend run again (into
                                    scratch, not the targets) counting the LDS-list records its
                                    quadrant waves walked; each step is one record for the
                                    wave's 64 pixels. Empty after other kinds of frame. */
} lrt_debug_buffer;
lrt_status lrt_raster_debug(lrt_context* ctx, int32_t which, void* host, size_t bytes, size_t* needed);

/* The kernel constant block one instance would be drawn with (host only):
 * FrameParams of lucabrtrender_amd/csrc/frame.h, the reference's setFrame +
 * setObject (modules/render/src/FrameParams.h:32-69, TileRasterizer.cpp:297-314). */
lrt_status lrt_frame_params(const lrt_projection* projection, const lrt_splat_instance* instance,
                            const lrt_render_settings* settings, uint32_t base, void* out, size_t bytes,
                            size_t* needed);

/* ---- the algorithms -------------------------------------------------------- */

/* gpu::PrefixSum::apply (modules/gpu/include/lrt/gpu/algo/PrefixSum.h:43-44):
 * out[i] = in[0] + ... + in[i-1] in u32, *total = the sum. Device pointers;
 * in and out must differ. Queued on the context's stream, not waited for. */
lrt_status lrt_prefix_sum(lrt_context* ctx, const uint32_t* in, uint32_t* out, uint32_t* total, uint32_t count);

/* The prefix sum of lrt_prefix_sum (same contract, same kernels) with no
 * allocation on any call and no context. No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). Its scratch
 * (look-back words and counters, about 768 KiB) is a static array in the
 * library's code object, one copy a device, held until the process ends.
 * count is at most LRT_PREFIX_SUM_STATIC_MAX (2^28; LRT_INVALID_ARGUMENT past
 * it). Runs on the current device, queued on `stream` (a hipStream_t; NULL is
 * the default stream), not waited for. Calls that overlap on one device share
 * that scratch and race: serialise them (one stream, or events). Messages:
 * lrt_last_error(NULL). */
#define LRT_PREFIX_SUM_STATIC_MAX (1u << 28)
lrt_status lrt_prefix_sum_static(const uint32_t* in, uint32_t* out, uint32_t* total, uint32_t count, void* stream);

/* gpu::RadixSort::sort (modules/gpu/include/lrt/gpu/algo/RadixSort.h:56-57):
 * stable sort of count (key, value) pairs by the key's low key_bits bits
 * (1..64), ceil(key_bits / 8) passes of at most eight bits. Keys are u64 here (the reference splits them in
 * keysLo/keysHi). scratch_* are same-sized buffers the passes alternate with;
 * the result is left in keys/values, or in scratch_* when *result_in_scratch is
 * set (an odd pass count; the reference copies back instead). Queued, not waited. */
lrt_status lrt_radix_sort(lrt_context* ctx, uint64_t* keys, uint32_t* values, uint64_t* scratch_keys,
                          uint32_t* scratch_values, uint32_t count, uint32_t key_bits, int32_t* result_in_scratch);

/* lrt_radix_sort with digits up to radix_bits (8..10) wide: ceil(key_bits /
 * radix_bits) passes, the same sorted result (no counterpart in the reference,
 * whose digits are 8 bits). */
lrt_status lrt_radix_sort_digits(lrt_context* ctx, uint64_t* keys, uint32_t* values, uint64_t* scratch_keys,
                                 uint32_t* scratch_values, uint32_t count, uint32_t key_bits, uint32_t radix_bits,
                                 int32_t* result_in_scratch);

/* Waits for everything queued on the context's stream. */
lrt_status lrt_context_synchronize(lrt_context* ctx);
/* Copies `bytes` of device memory to the host after everything queued on the
 * context's stream: how a caller reads buffers the library owns (the levels of
 * an lrt_lod_cloud, the clouds of a cut). */
lrt_status lrt_copy_to_host(lrt_context* ctx, const void* device, void* host, size_t bytes);

/* ---- images ------------------------------------------------------------------
 * io::writeExr (modules/io/src/Exr.cpp:130-157): an OpenEXR scanline file,
 * ZIP-compressed, channels A B G R (half when half != 0, else float) and Z
 * (float, when depth is not NULL), rows top first from images stored bottom
 * row first as the rasteriser stores them. rgba (width * height * 4) and depth
 * (width * height) are HOST memory. LRT_IO_FAILURE when the file cannot be
 * written; the message: lrt_last_error(NULL). */
lrt_status lrt_write_exr(const char* path, uint32_t width, uint32_t height, const float* rgba, const float* depth,
                         int32_t half);

/* ---- splat files --------------------------------------------------------------
 * io::RawSplats and io::SplatEncoding
 * (modules/io/include/lrt/io/RawSplats.h:15-77): a file as one float a property
 * a record, in the file's order, and where each field sits. Modes are numbered as RawSplats.h:42-58 numbers them:
 *   opacity_mode   0 logit, 1 linear, 2 byte (v / 255)
 *   scale_mode     0 log, 1 linear (log(max(v, 1e-30))), 2 SPZ byte (v / 16 - 10)
 *   colour_mode    0 SH DC (0.5 + C0 v), 1 linear, 2 byte (v / 255),
 *                  3 SPZ byte (0.5 + C0 ((v / 255 - 0.5) / 0.15))
 *   rotation_mode  0 float (w, x, y, z), 1 byte ((v - 128) / 128 each),
 *                  2 SPZ v2 (x y z bytes v / 127.5 - 1 at rot_x..rot_z, w from unit length),
 *                  3 SPZ v3 smallest three (a u32: low 16 bits at rot_x, high 16 at rot_y)
 *   rest_mode      0 float, 1 byte ((v - 128) / 128)
 * rest_colour_outer 1: rest coefficients rrr.. ggg.. bbb.. with
 * file_per_colour a colour (a 3DGS PLY); 0: rgb per basis. Positions are
 * multiplied by 2^-fractional_bits (SPZ's fixed point); flip_yz 1: the file
 * is right-up-back (SPZ), read right-down-front (y and z of positions and
 * rotations, and the rest coefficients odd in y xor z, change sign). */
typedef struct lrt_raw_splats lrt_raw_splats;
typedef struct {
    uint32_t floats_per_record;
    uint32_t x, y, z, opacity, scale0, scale1, scale2, rot_w, rot_x, rot_y, rot_z, dc0, dc1, dc2;
    uint32_t rest_base, rest_per_colour, file_per_colour, rest_colour_outer;
    uint32_t opacity_mode, scale_mode, colour_mode, rotation_mode, rest_mode;
    uint32_t fractional_bits, flip_yz;
} lrt_splat_encoding;

/* io::readSplats (modules/io/src/SplatReaders.cpp:32-328) by extension: a
 * binary little-endian 3DGS PLY, a .splat file or an SPZ file of version 2 or
 * 3 (one gzip stream), read on the host. SPZ version 4 (zstd streams), SPZ
 * version 1 (float16 positions) and .sog are refused (LRT_UNSUPPORTED): this
 * build has neither zstd nor an image decoder; an SPZ whose header or streams
 * do not hold together is LRT_IO_FAILURE. The reference's codes and messages
 * otherwise: not a PLY, no
 * vertices, a missing 3DGS property -> LRT_INVALID_ARGUMENT; ASCII or
 * big-endian, a list in the vertex element, harmonics not contiguous or not a
 * whole degree -> LRT_UNSUPPORTED; a file shorter than its vertices ->
 * LRT_IO_FAILURE. The message: lrt_last_error(NULL). */
lrt_status lrt_splats_read(const char* path, lrt_raw_splats** out);
/* Raw records from memory (another reader's output): count * floats_per_record floats, copied. */
lrt_status lrt_raw_splats_create(const float* records, uint32_t count, const lrt_splat_encoding* encoding,
                                 lrt_raw_splats** out);
void lrt_raw_splats_destroy(lrt_raw_splats* raw);
uint32_t lrt_raw_splats_count(const lrt_raw_splats* raw);
lrt_status lrt_raw_splats_encoding(const lrt_raw_splats* raw, lrt_splat_encoding* out);
/* The records, host memory owned by raw: count * floats_per_record floats. */
const float* lrt_raw_splats_records(const lrt_raw_splats* raw);

/* What a decode made: records declared and kept, the rest basis functions kept
 * a colour, the packed row's words, and the kept positions' box
 * (scene::GpuSplats count/declared/bounds, GpuClouds.h:41-76); and the GPU
 * milliseconds of the records' copies to the device and of the decode's kernels. */
typedef struct {
    uint32_t declared, count, rest_per_colour, sh_words;
    float bounds_min[3], bounds_max[3];
    float upload_ms, decode_ms;
} lrt_cloud_info;

/* The sh_words a decode of raw at max_degree packs a splat. */
uint32_t lrt_cloud_sh_words(const lrt_raw_splats* raw, uint32_t max_degree);

/* scene::CloudLoader::upload (modules/scene/src/GpuClouds.cpp:151-227) with
 * splat_validate.slang and splat_decode.slang: records kept when the position
 * and log scales are finite, |log scale| < 60 and opacity >= 1/255, packed in
 * file order into the lrt_splats layout at caller device buffers sized for the
 * declared count (positions 16 B, shape 16 B, sh 4 * sh_words B a splat; sh
 * may be NULL when no rest coefficient is kept). Rest coefficients are kept up
 * to min(the file's degree, max_degree). LRT_INVALID_ARGUMENT when no record
 * survives. Synchronous; the context's last frame's debug buffers are lost. */
lrt_status lrt_cloud_decode(lrt_context* ctx, const lrt_raw_splats* raw, uint32_t max_degree, float* positions,
                            uint32_t* shape, uint32_t* sh, lrt_cloud_info* info);

#ifdef __cplusplus
}
#endif
#endif
