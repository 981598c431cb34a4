// The extern "C" boundary of include/lrt_hip.h over the C++20 rasteriser.
#include <hip/hip_runtime.h>

#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/lrt_hip.h"
#include "exr.h"
#include "lod.h"
#include "rasterizer.h"

struct lrt_context {
    lrt::TileRasterizer raster;
    lrt::lod::CutSelector cut;
    std::vector<lrt_splats> cutClouds;   // what the last lrt_lod_select's instances point at
    std::string error;
};

struct lrt_lod_cloud {
    lrt::lod::LodCloud lod;
};

struct lrt_raw_splats {
    lrt::io::RawSplats raw;
};

namespace {

// The message of the last failure without a context (lrt_context_create).
thread_local std::string gNoContextError;

lrt_status finish(lrt_context* ctx, const lrt::Status& s) {
    if (ctx != nullptr) ctx->error = s.ok() ? std::string() : s.message;
    return static_cast<lrt_status>(static_cast<int>(s.code));
}

lrt::Mat4 toMat(const double m[16]) {
    lrt::Mat4 out;
    for (int k = 0; k < 16; ++k) out.m[static_cast<size_t>(k)] = m[k];
    return out;
}

lrt::Projection toProjection(const lrt_projection& p) {
    lrt::Projection out;
    out.worldToView = toMat(p.world_to_view);
    out.focalX = p.focal_x;
    out.focalY = p.focal_y;
    out.centreX = p.centre_x;
    out.centreY = p.centre_y;
    out.nearZ = p.near_z;
    out.farZ = p.far_z;
    out.orthographic = p.orthographic != 0;
    out.eyeWorld = {p.eye_world[0], p.eye_world[1], p.eye_world[2]};
    return out;
}

lrt::RenderSettings toSettings(const lrt_render_settings& s) {
    lrt::RenderSettings out;
    out.width = s.width;
    out.height = s.height;
    out.antialias = s.antialias != 0;
    out.maxShDegree = s.max_sh_degree;
    for (int k = 0; k < 4; ++k) out.background[static_cast<size_t>(k)] = s.background[k];
    out.depthMode = s.depth_mode != 0 ? 1u : 0u;
    out.depthThreshold = s.depth_threshold;
    out.linearise = s.linearise != 0;
    out.timeStages = s.time_stages != 0;
    out.dilation = s.dilation;
    out.filter3d = s.filter3d != 0;
    out.blendBatch = s.blend_batch == 0 ? 256u : s.blend_batch;
    out.blendSegment = s.blend_segment;
    out.keepWalkState = s.keep_walk_state != 0;
    out.heaviestTilesFirst = s.tile_order == 1;
    out.persistentBlend = s.persistent_blend != 0;
    out.useGraph = s.use_graph != 0;
    out.tileSize = s.tile_size == 0 ? 16u : s.tile_size;
    out.radixBits = s.radix_bits;
    return out;
}

lrt::SplatEditParams toEdit(const lrt_splat_edit& e) {
    lrt::SplatEditParams out;
    out.active = e.active != 0 ? 1u : 0u;
    out.shape = e.shape;
    out.mode = e.mode;
    out.invert = e.invert != 0 ? 1u : 0u;
    for (int k = 0; k < 3; ++k) {
        out.centre[k] = e.centre[k];
        out.size[k] = e.size[k];
        out.tint[k] = e.tint[k];
    }
    out.saturation = e.saturation;
    out.brightness = e.brightness;
    out.opacity = e.opacity;
    out.minOpacity = e.min_opacity;
    out.maxScale = e.max_scale;
    return out;
}

lrt::SplatCloud toCloud(const lrt_splats& c) {
    lrt::SplatCloud out;
    out.positions = reinterpret_cast<const float4*>(c.positions);
    out.shape = reinterpret_cast<const uint4*>(c.shape);
    out.sh = c.sh;
    out.filter3d = c.filter3d;
    out.count = c.count;
    out.restPerColour = c.rest_per_colour;
    out.shWords = c.sh_words;
    return out;
}

}   // namespace

namespace {

lrt::Status toBatch(const lrt_view* views, uint32_t view_count, const lrt_splat_instance* instances,
                    uint32_t instance_count, std::vector<lrt::View>& outViews, std::vector<lrt::SplatCloud>& clouds,
                    std::vector<lrt::SplatInstance>& list) {
    if ((view_count > 0 && views == nullptr) || (instance_count > 0 && instances == nullptr)) {
        return lrt::Status::error(lrt::ErrorCode::InvalidArgument, "null argument");
    }
    outViews.resize(view_count);
    for (uint32_t v = 0; v < view_count; ++v) {
        outViews[v].projection = toProjection(views[v].projection);
        outViews[v].targets.width = views[v].targets.width;
        outViews[v].targets.height = views[v].targets.height;
        outViews[v].targets.colour = reinterpret_cast<float4*>(views[v].targets.colour);
        outViews[v].targets.depth = views[v].targets.depth;
    }
    clouds.resize(instance_count);
    list.resize(instance_count);
    for (uint32_t k = 0; k < instance_count; ++k) {
        if (instances[k].splats != nullptr) {
            clouds[k] = toCloud(*instances[k].splats);
            list[k].splats = &clouds[k];
        }
        list[k].objectToWorld = toMat(instances[k].object_to_world);
        list[k].edit = toEdit(instances[k].edit);
    }
    return {};
}

void toStats(const lrt::FrameStats& fs, lrt_frame_stats* stats) {
    if (stats == nullptr) return;
    stats->splats = fs.splats;
    stats->visible = fs.visible;
    stats->pairs = fs.pairs;
    stats->project_ms = fs.projectMs;
    stats->depth_sort_ms = fs.depthSortMs;
    stats->counts_ms = fs.countsMs;
    stats->emit_ms = fs.emitMs;
    stats->tile_sort_ms = fs.tileSortMs;
    stats->blend_ms = fs.blendMs;
    stats->total_ms = fs.totalMs;
    stats->key_bits = fs.keyBits;
    stats->sort_passes = fs.sortPasses;
}

lrt_splat_encoding toEncodingC(const lrt::io::SplatEncoding& e) {
    lrt_splat_encoding c{};
    c.floats_per_record = e.floatsPerRecord;
    c.x = e.x; c.y = e.y; c.z = e.z; c.opacity = e.opacity;
    c.scale0 = e.scale0; c.scale1 = e.scale1; c.scale2 = e.scale2;
    c.rot_w = e.rotW; c.rot_x = e.rotX; c.rot_y = e.rotY; c.rot_z = e.rotZ;
    c.dc0 = e.dc0; c.dc1 = e.dc1; c.dc2 = e.dc2;
    c.rest_base = e.restBase; c.rest_per_colour = e.restPerColour; c.file_per_colour = e.filePerColour;
    c.rest_colour_outer = e.restColourOuter;
    c.opacity_mode = static_cast<uint32_t>(e.opacityMode);
    c.scale_mode = static_cast<uint32_t>(e.scaleMode);
    c.colour_mode = static_cast<uint32_t>(e.colourMode);
    c.rotation_mode = static_cast<uint32_t>(e.rotationMode);
    c.rest_mode = static_cast<uint32_t>(e.restMode);
    c.fractional_bits = e.fractionalBits;
    c.flip_yz = e.flipYZ;
    return c;
}

// The C encoding checked and converted: modes in range, rest counts one of 0, 3, 8, 15.
lrt::Status fromEncodingC(const lrt_splat_encoding& c, lrt::io::SplatEncoding& e) {
    using lrt::ErrorCode;
    using lrt::Status;
    using E = lrt::io::SplatEncoding;
    if (c.opacity_mode > 2 || c.scale_mode > 2 || c.colour_mode > 3 || c.rotation_mode > 3 || c.rest_mode > 1) {
        return Status::error(ErrorCode::InvalidArgument, "an encoding mode this build does not decode");
    }
    if (c.fractional_bits > 24) return Status::error(ErrorCode::InvalidArgument, "fractional_bits above 24");
    if (c.rest_per_colour != 0 && c.rest_per_colour != 3 && c.rest_per_colour != 8 && c.rest_per_colour != 15) {
        return Status::error(ErrorCode::InvalidArgument, "rest_per_colour must be 0, 3, 8 or 15");
    }
    e.floatsPerRecord = c.floats_per_record;
    e.x = c.x; e.y = c.y; e.z = c.z; e.opacity = c.opacity;
    e.scale0 = c.scale0; e.scale1 = c.scale1; e.scale2 = c.scale2;
    e.rotW = c.rot_w; e.rotX = c.rot_x; e.rotY = c.rot_y; e.rotZ = c.rot_z;
    e.dc0 = c.dc0; e.dc1 = c.dc1; e.dc2 = c.dc2;
    e.restBase = c.rest_base; e.restPerColour = c.rest_per_colour; e.filePerColour = c.file_per_colour;
    e.restColourOuter = c.rest_colour_outer != 0 ? 1u : 0u;
    e.opacityMode = static_cast<E::Opacity>(c.opacity_mode);
    e.scaleMode = static_cast<E::Scale>(c.scale_mode);
    e.colourMode = static_cast<E::Colour>(c.colour_mode);
    e.rotationMode = static_cast<E::Rotation>(c.rotation_mode);
    e.restMode = static_cast<E::Rest>(c.rest_mode);
    e.fractionalBits = c.fractional_bits;
    e.flipYZ = c.flip_yz != 0 ? 1u : 0u;
    return {};
}

lrt_status noContext(const lrt::Status& s) {
    gNoContextError = s.ok() ? std::string() : s.message;
    return static_cast<lrt_status>(static_cast<int>(s.code));
}

}   // namespace

extern "C" {

void lrt_lens_default(lrt_lens* out) {
    if (out == nullptr) return;
    const lrt::Lens l;
    out->orthographic = l.orthographic ? 1 : 0;
    out->focal = l.focal;
    out->haperture = l.haperture;
    out->vaperture = l.vaperture;
    out->near_z = l.nearZ;
    out->far_z = l.farZ;
    out->window_translate[0] = l.windowTranslate[0];
    out->window_translate[1] = l.windowTranslate[1];
    out->window_scale[0] = l.windowScale[0];
    out->window_scale[1] = l.windowScale[1];
    out->window_roll = l.windowRoll;
}

void lrt_camera_looking_at(const double eye[3], const double target[3], const double up[3], lrt_camera* out) {
    if (out == nullptr || eye == nullptr || target == nullptr) return;
    const lrt::Vec3 u = up != nullptr ? lrt::Vec3{up[0], up[1], up[2]} : lrt::Vec3{0.0, 1.0, 0.0};
    const lrt::Camera c = lrt::Camera::lookingAt({eye[0], eye[1], eye[2]}, {target[0], target[1], target[2]}, u);
    for (int k = 0; k < 16; ++k) out->camera_to_world[k] = c.cameraToWorld.m[static_cast<size_t>(k)];
    lrt_lens_default(&out->lens);
}

lrt_status lrt_projection_for(const lrt_camera* camera, uint32_t width, uint32_t height, lrt_projection* out) {
    if (camera == nullptr || out == nullptr) return LRT_INVALID_ARGUMENT;
    lrt::Camera c;
    c.cameraToWorld = toMat(camera->camera_to_world);
    c.lens.orthographic = camera->lens.orthographic != 0;
    c.lens.focal = camera->lens.focal;
    c.lens.haperture = camera->lens.haperture;
    c.lens.vaperture = camera->lens.vaperture;
    c.lens.nearZ = camera->lens.near_z;
    c.lens.farZ = camera->lens.far_z;
    for (int k = 0; k < 2; ++k) {
        c.lens.windowTranslate[k] = camera->lens.window_translate[k];
        c.lens.windowScale[k] = camera->lens.window_scale[k];
    }
    c.lens.windowRoll = camera->lens.window_roll;
    const lrt::Projection p = lrt::projectionFor(c, width, height);
    for (int k = 0; k < 16; ++k) out->world_to_view[k] = p.worldToView.m[static_cast<size_t>(k)];
    out->focal_x = p.focalX;
    out->focal_y = p.focalY;
    out->centre_x = p.centreX;
    out->centre_y = p.centreY;
    out->near_z = p.nearZ;
    out->far_z = p.farZ;
    out->orthographic = p.orthographic ? 1 : 0;
    out->eye_world[0] = p.eyeWorld.x;
    out->eye_world[1] = p.eyeWorld.y;
    out->eye_world[2] = p.eyeWorld.z;
    return LRT_OK;
}

void lrt_render_settings_default(lrt_render_settings* out) {
    if (out == nullptr) return;
    const lrt::RenderSettings s;
    std::memset(out, 0, sizeof *out);
    out->width = s.width;
    out->height = s.height;
    out->antialias = s.antialias ? 1 : 0;
    out->max_sh_degree = s.maxShDegree;
    for (int k = 0; k < 4; ++k) out->background[k] = s.background[static_cast<size_t>(k)];
    out->depth_mode = 0;
    out->depth_threshold = s.depthThreshold;
    out->linearise = s.linearise ? 1 : 0;
    out->time_stages = 0;
    out->dilation = s.dilation;
    out->filter3d = 0;
    out->blend_batch = 256;
    out->blend_segment = s.blendSegment;
    out->keep_walk_state = s.keepWalkState ? 1 : 0;
    out->tile_order = 1;
    out->persistent_blend = 0;
    out->use_graph = 0;
    out->tile_size = s.tileSize;
    out->radix_bits = s.radixBits;
}

void lrt_splat_edit_default(lrt_splat_edit* out) {
    if (out == nullptr) return;
    const lrt::SplatEditParams e;
    std::memset(out, 0, sizeof *out);
    out->active = 0;
    out->shape = e.shape;
    out->mode = e.mode;
    out->invert = 0;
    for (int k = 0; k < 3; ++k) {
        out->centre[k] = e.centre[k];
        out->size[k] = e.size[k];
        out->tint[k] = e.tint[k];
    }
    out->saturation = e.saturation;
    out->brightness = e.brightness;
    out->opacity = e.opacity;
    out->min_opacity = e.minOpacity;
    out->max_scale = e.maxScale;
}

lrt_status lrt_context_create(int device, void* stream, lrt_context** out) {
    if (out == nullptr) return LRT_INVALID_ARGUMENT;
    *out = nullptr;
    auto* ctx = new (std::nothrow) lrt_context();
    if (ctx == nullptr) return LRT_OUT_OF_MEMORY;
    const lrt::Status s = ctx->raster.create(device, static_cast<hipStream_t>(stream));
    if (!s.ok()) {
        delete ctx;
        return noContext(s);
    }
    *out = ctx;
    return noContext({});
}

void lrt_context_destroy(lrt_context* ctx) { delete ctx; }

lrt_status lrt_context_set_stream(lrt_context* ctx, void* stream) {
    if (ctx == nullptr) return LRT_INVALID_ARGUMENT;
    ctx->raster.setStream(static_cast<hipStream_t>(stream));
    return finish(ctx, {});
}

const char* lrt_last_error(const lrt_context* ctx) {
    return ctx != nullptr ? ctx->error.c_str() : gNoContextError.c_str();
}

namespace {

lrt_status renderFrame(lrt_context* ctx, const lrt_projection* projection, const lrt_splat_instance* instances,
                       uint32_t instance_count, const lrt_point_instance* points, uint32_t point_count,
                       const lrt_render_targets* under, const lrt_render_settings* settings,
                       lrt_render_targets* targets, lrt_frame_stats* stats) {
    if (ctx == nullptr) return LRT_INVALID_ARGUMENT;
    if (projection == nullptr || settings == nullptr || targets == nullptr || (instance_count > 0 && instances == nullptr) ||
        (point_count > 0 && points == nullptr)) {
        return finish(ctx, lrt::Status::error(lrt::ErrorCode::InvalidArgument, "null argument"));
    }
    std::vector<lrt::SplatCloud> clouds(instance_count);
    std::vector<lrt::SplatInstance> list(instance_count);
    for (uint32_t k = 0; k < instance_count; ++k) {
        if (instances[k].splats != nullptr) {
            clouds[k] = toCloud(*instances[k].splats);
            list[k].splats = &clouds[k];
        }
        list[k].objectToWorld = toMat(instances[k].object_to_world);
        list[k].edit = toEdit(instances[k].edit);
    }
    std::vector<lrt::PointCloud> pointClouds(point_count);
    std::vector<lrt::PointInstance> pointList(point_count);
    for (uint32_t k = 0; k < point_count; ++k) {
        const lrt_point_instance& in = points[k];
        if (in.points != nullptr) {
            pointClouds[k].positions = reinterpret_cast<const float4*>(in.points->positions);
            pointClouds[k].colours = in.points->colours;
            pointClouds[k].count = in.points->count;
            pointList[k].points = &pointClouds[k];
        }
        pointList[k].objectToWorld = toMat(in.object_to_world);
        if (in.style.size_mode > 1) {
            return finish(ctx, lrt::Status::error(lrt::ErrorCode::InvalidArgument, "size_mode must be 0 or 1"));
        }
        pointList[k].style.sizeMode = in.style.size_mode == 1 ? lrt::PointStyle::Size::Pixels
                                                              : lrt::PointStyle::Size::World;
        pointList[k].style.size = in.style.size;
        pointList[k].style.constantColour = in.style.constant_colour != 0;
        for (int c = 0; c < 3; ++c) pointList[k].style.colour[static_cast<size_t>(c)] = in.style.colour[c];
    }
    lrt::RenderTargets t;
    t.width = targets->width;
    t.height = targets->height;
    t.colour = reinterpret_cast<float4*>(targets->colour);
    t.depth = targets->depth;
    lrt::RenderTargets u;
    if (under != nullptr) {
        u.width = under->width;
        u.height = under->height;
        u.colour = reinterpret_cast<float4*>(under->colour);
        u.depth = under->depth;
    }
    lrt::FrameStats fs;
    const lrt::Status s = ctx->raster.render(toProjection(*projection), list, toSettings(*settings), t, &fs, pointList,
                                             under != nullptr ? &u : nullptr);
    if (s.ok() && stats != nullptr) {
        stats->splats = fs.splats;
        stats->visible = fs.visible;
        stats->pairs = fs.pairs;
        stats->project_ms = fs.projectMs;
        stats->depth_sort_ms = fs.depthSortMs;
        stats->counts_ms = fs.countsMs;
        stats->emit_ms = fs.emitMs;
        stats->tile_sort_ms = fs.tileSortMs;
        stats->blend_ms = fs.blendMs;
        stats->total_ms = fs.totalMs;
        stats->key_bits = fs.keyBits;
        stats->sort_passes = fs.sortPasses;
    }
    return finish(ctx, s);
}

}   // namespace

lrt_status lrt_raster_render(lrt_context* ctx, const lrt_projection* projection, const lrt_splat_instance* instances,
                             uint32_t instance_count, const lrt_render_settings* settings, lrt_render_targets* targets,
                             lrt_frame_stats* stats) {
    return renderFrame(ctx, projection, instances, instance_count, nullptr, 0, nullptr, settings, targets, stats);
}

lrt_status lrt_raster_render_composite(lrt_context* ctx, const lrt_projection* projection,
                                       const lrt_splat_instance* instances, uint32_t instance_count,
                                       const lrt_point_instance* points, uint32_t point_count,
                                       const lrt_render_targets* under, const lrt_render_settings* settings,
                                       lrt_render_targets* targets, lrt_frame_stats* stats) {
    return renderFrame(ctx, projection, instances, instance_count, points, point_count, under, settings, targets,
                       stats);
}

void lrt_point_style_default(lrt_point_style* out) {
    if (out == nullptr) return;
    const lrt::PointStyle st;
    std::memset(out, 0, sizeof *out);
    out->size_mode = st.sizeMode == lrt::PointStyle::Size::Pixels ? 1u : 0u;
    out->size = st.size;
    out->constant_colour = st.constantColour ? 1 : 0;
    for (int c = 0; c < 3; ++c) out->colour[c] = st.colour[static_cast<size_t>(c)];
}


lrt_status lrt_reference_render(lrt_context* ctx, const lrt_projection* projection,
                                const lrt_splat_instance* instances, uint32_t instance_count,
                                const lrt_render_settings* settings, double* colour, double* depth,
                                int32_t keyed) {
    if (ctx == nullptr) return LRT_INVALID_ARGUMENT;
    if (projection == nullptr || settings == nullptr || (instance_count > 0 && instances == nullptr)) {
        return finish(ctx, lrt::Status::error(lrt::ErrorCode::InvalidArgument, "null argument"));
    }
    std::vector<lrt::View> none;
    std::vector<lrt::SplatCloud> clouds;
    std::vector<lrt::SplatInstance> inst;
    lrt::Status s = toBatch(nullptr, 0, instances, instance_count, none, clouds, inst);
    if (!s.ok()) return finish(ctx, s);
    return finish(ctx, ctx->raster.referenceRender(toProjection(*projection), inst, toSettings(*settings), colour,
                                                   depth, keyed != 0));
}

lrt_status lrt_reference_render_pixels(lrt_context* ctx, const lrt_projection* projection,
                                       const lrt_splat_instance* instances, uint32_t instance_count,
                                       const lrt_render_settings* settings, const uint32_t* pixels,
                                       uint32_t pixel_count, double* colour, double* depth, int32_t keyed) {
    if (ctx == nullptr) return LRT_INVALID_ARGUMENT;
    if (projection == nullptr || settings == nullptr || (instance_count > 0 && instances == nullptr) ||
        (pixel_count > 0 && pixels == nullptr)) {
        return finish(ctx, lrt::Status::error(lrt::ErrorCode::InvalidArgument, "null argument"));
    }
    if (pixel_count == 0) return finish(ctx, {});
    std::vector<lrt::View> none;
    std::vector<lrt::SplatCloud> clouds;
    std::vector<lrt::SplatInstance> inst;
    lrt::Status s = toBatch(nullptr, 0, instances, instance_count, none, clouds, inst);
    if (!s.ok()) return finish(ctx, s);
    return finish(ctx, ctx->raster.referenceRender(toProjection(*projection), inst, toSettings(*settings), colour,
                                                   depth, keyed != 0, pixels, pixel_count));
}

lrt_status lrt_raster_render_views(lrt_context* ctx, const lrt_view* views, uint32_t view_count,
                                   const lrt_splat_instance* instances, uint32_t instance_count,
                                   const lrt_render_settings* settings, lrt_frame_stats* stats) {
    if (ctx == nullptr) return LRT_INVALID_ARGUMENT;
    if (settings == nullptr) return finish(ctx, lrt::Status::error(lrt::ErrorCode::InvalidArgument, "null argument"));
    std::vector<lrt::View> list;
    std::vector<lrt::SplatCloud> clouds;
    std::vector<lrt::SplatInstance> inst;
    lrt::Status s = toBatch(views, view_count, instances, instance_count, list, clouds, inst);
    if (!s.ok()) return finish(ctx, s);
    lrt::FrameStats fs;
    s = ctx->raster.renderViews(list, inst, toSettings(*settings), &fs);
    if (s.ok()) toStats(fs, stats);
    return finish(ctx, s);
}

lrt_status lrt_autotune_views(lrt_context* ctx, const lrt_view* views, uint32_t view_count,
                              const lrt_splat_instance* instances, uint32_t instance_count,
                              lrt_render_settings* settings, float* ms) {
    if (ctx == nullptr) return LRT_INVALID_ARGUMENT;
    if (settings == nullptr) return finish(ctx, lrt::Status::error(lrt::ErrorCode::InvalidArgument, "null argument"));
    std::vector<lrt::View> list;
    std::vector<lrt::SplatCloud> clouds;
    std::vector<lrt::SplatInstance> inst;
    lrt::Status s = toBatch(views, view_count, instances, instance_count, list, clouds, inst);
    if (!s.ok()) return finish(ctx, s);
    lrt::RenderSettings best = toSettings(*settings);
    float bestMs = 0.0F;
    s = ctx->raster.autotuneViews(list, inst, best, &bestMs);
    if (!s.ok()) return finish(ctx, s);
    settings->blend_batch = best.blendBatch;
    settings->persistent_blend = best.persistentBlend ? 1 : 0;
    settings->use_graph = best.useGraph ? 1 : 0;
    settings->tile_size = best.tileSize;
    settings->radix_bits = best.radixBits;
    if (ms != nullptr) *ms = bestMs;
    return finish(ctx, s);
}

lrt_status lrt_raster_render_diff(lrt_context* ctx, const lrt_projection* projection,
                                  const lrt_diff_splats* splats, const double object_to_world[16],
                                  const lrt_render_settings* settings, lrt_render_targets* targets,
                                  lrt_frame_stats* stats) {
    if (ctx == nullptr) return LRT_INVALID_ARGUMENT;
    if (projection == nullptr || splats == nullptr || settings == nullptr || targets == nullptr) {
        return finish(ctx, lrt::Status::error(lrt::ErrorCode::InvalidArgument, "null argument"));
    }
    lrt::kern::DiffCloud cloud{splats->means, splats->log_scales, splats->quats, splats->opacity_logits, splats->sh,
                               splats->filter3d, splats->count, splats->rest_per_colour};
    lrt::Mat4 o2w;
    if (object_to_world != nullptr) o2w = toMat(object_to_world);
    lrt::RenderTargets t;
    t.width = targets->width;
    t.height = targets->height;
    t.colour = reinterpret_cast<float4*>(targets->colour);
    t.depth = targets->depth;
    lrt::FrameStats fs;
    const lrt::Status s = ctx->raster.renderDiff(toProjection(*projection), cloud, o2w, toSettings(*settings), t, &fs);
    if (s.ok() && stats != nullptr) {
        stats->splats = fs.splats;
        stats->visible = fs.visible;
        stats->pairs = fs.pairs;
        stats->project_ms = fs.projectMs;
        stats->depth_sort_ms = fs.depthSortMs;
        stats->counts_ms = fs.countsMs;
        stats->emit_ms = fs.emitMs;
        stats->tile_sort_ms = fs.tileSortMs;
        stats->blend_ms = fs.blendMs;
        stats->total_ms = fs.totalMs;
        stats->key_bits = fs.keyBits;
        stats->sort_passes = fs.sortPasses;
    }
    return finish(ctx, s);
}

lrt_status lrt_raster_backward(lrt_context* ctx, const float* grad_colour, const lrt_diff_grads* grads,
                               float* millis) {
    if (ctx == nullptr) return LRT_INVALID_ARGUMENT;
    if (grads == nullptr) return finish(ctx, lrt::Status::error(lrt::ErrorCode::InvalidArgument, "null argument"));
    const lrt::kern::DiffGrads g{grads->means, grads->log_scales, grads->quats, grads->opacity_logits, grads->sh};
    return finish(ctx, ctx->raster.backward(reinterpret_cast<const float4*>(grad_colour), g, millis));
}

lrt_status lrt_raster_render_diff_views(lrt_context* ctx, const lrt_view* views, uint32_t view_count,
                                        const lrt_diff_splats* splats, const double object_to_world[16],
                                        const lrt_render_settings* settings, lrt_frame_stats* stats) {
    if (ctx == nullptr) return LRT_INVALID_ARGUMENT;
    if (splats == nullptr || settings == nullptr) {
        return finish(ctx, lrt::Status::error(lrt::ErrorCode::InvalidArgument, "null argument"));
    }
    std::vector<lrt::View> list;
    std::vector<lrt::SplatCloud> clouds;
    std::vector<lrt::SplatInstance> inst;
    lrt::Status s = toBatch(views, view_count, nullptr, 0, list, clouds, inst);
    if (!s.ok()) return finish(ctx, s);
    const lrt::kern::DiffCloud cloud{splats->means, splats->log_scales, splats->quats, splats->opacity_logits,
                                     splats->sh, splats->filter3d, splats->count, splats->rest_per_colour};
    lrt::Mat4 o2w;
    if (object_to_world != nullptr) o2w = toMat(object_to_world);
    lrt::FrameStats fs;
    s = ctx->raster.renderDiffViews(list, cloud, o2w, toSettings(*settings), &fs);
    if (s.ok()) toStats(fs, stats);
    return finish(ctx, s);
}

lrt_status lrt_raster_backward_views(lrt_context* ctx, const float* const* grad_colours, uint32_t view_count,
                                     const lrt_diff_grads* grads, float* millis) {
    if (ctx == nullptr) return LRT_INVALID_ARGUMENT;
    if (grads == nullptr || (view_count > 0 && grad_colours == nullptr)) {
        return finish(ctx, lrt::Status::error(lrt::ErrorCode::InvalidArgument, "null argument"));
    }
    std::vector<const float4*> list(view_count);
    for (uint32_t v = 0; v < view_count; ++v) list[v] = reinterpret_cast<const float4*>(grad_colours[v]);
    const lrt::kern::DiffGrads g{grads->means, grads->log_scales, grads->quats, grads->opacity_logits, grads->sh};
    return finish(ctx, ctx->raster.backwardViews(list, g, millis));
}

lrt_status lrt_raster_debug(lrt_context* ctx, int32_t which, void* host, size_t bytes, size_t* needed) {
    if (ctx == nullptr) return LRT_INVALID_ARGUMENT;
    return finish(ctx, ctx->raster.debugCopy(which, host, bytes, needed));
}

lrt_status lrt_frame_params(const lrt_projection* projection, const lrt_splat_instance* instance,
                            const lrt_render_settings* settings, uint32_t base, void* out, size_t bytes,
                            size_t* needed) {
    if (needed != nullptr) *needed = sizeof(lrt::FrameParams);
    if (projection == nullptr || instance == nullptr || settings == nullptr) return LRT_INVALID_ARGUMENT;
    if (out == nullptr || bytes == 0) return LRT_OK;
    if (bytes < sizeof(lrt::FrameParams)) return LRT_INVALID_ARGUMENT;
    lrt::SplatCloud cloud;
    lrt::SplatInstance inst;
    if (instance->splats != nullptr) {
        cloud = toCloud(*instance->splats);
        inst.splats = &cloud;
    }
    inst.objectToWorld = toMat(instance->object_to_world);
    inst.edit = toEdit(instance->edit);
    const lrt::FrameParams f = lrt::frameParamsFor(toProjection(*projection), inst, toSettings(*settings), base);
    std::memcpy(out, &f, sizeof f);
    return LRT_OK;
}

lrt_status lrt_prefix_sum(lrt_context* ctx, const uint32_t* in, uint32_t* out, uint32_t* total, uint32_t count) {
    if (ctx == nullptr) return LRT_INVALID_ARGUMENT;
    return finish(ctx, ctx->raster.prefixSum(in, out, total, count));
}

lrt_status lrt_prefix_sum_static(const uint32_t* in, uint32_t* out, uint32_t* total, uint32_t count, void* stream) {
    static_assert(LRT_PREFIX_SUM_STATIC_MAX == lrt::kern::kStaticScanMax, "the header's limit is the scan's");
    if (in == nullptr || out == nullptr || total == nullptr) {
        return noContext(lrt::Status::error(lrt::ErrorCode::InvalidArgument, "null buffer"));
    }
    if (in == out) {
        return noContext(lrt::Status::error(lrt::ErrorCode::InvalidArgument, "the scan's input and output must differ"));
    }
    if (count > LRT_PREFIX_SUM_STATIC_MAX) {
        return noContext(lrt::Status::error(lrt::ErrorCode::InvalidArgument,
                                            "count " + std::to_string(count) + " exceeds the static scan's maximum " +
                                                std::to_string(LRT_PREFIX_SUM_STATIC_MAX)));
    }
    const hipError_t e = lrt::kern::scanExclusiveStatic(in, out, count, total, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) {
        return noContext(lrt::Status::error(lrt::ErrorCode::DeviceFailure,
                                            std::string("static prefix sum: ") + hipGetErrorString(e)));
    }
    return noContext({});
}

lrt_status lrt_radix_sort(lrt_context* ctx, uint64_t* keys, uint32_t* values, uint64_t* scratch_keys,
                          uint32_t* scratch_values, uint32_t count, uint32_t key_bits, int32_t* result_in_scratch) {
    if (ctx == nullptr) return LRT_INVALID_ARGUMENT;
    bool alt = false;
    const lrt::Status s =
        ctx->raster.radixSort(keys, values, scratch_keys, scratch_values, count, key_bits, &alt);
    if (result_in_scratch != nullptr) *result_in_scratch = alt ? 1 : 0;
    return finish(ctx, s);
}

lrt_status lrt_radix_sort_digits(lrt_context* ctx, uint64_t* keys, uint32_t* values, uint64_t* scratch_keys,
                                 uint32_t* scratch_values, uint32_t count, uint32_t key_bits, uint32_t radix_bits,
                                 int32_t* result_in_scratch) {
    if (ctx == nullptr) return LRT_INVALID_ARGUMENT;
    bool alt = false;
    const lrt::Status s =
        ctx->raster.radixSort(keys, values, scratch_keys, scratch_values, count, key_bits, &alt, radix_bits);
    if (result_in_scratch != nullptr) *result_in_scratch = alt ? 1 : 0;
    return finish(ctx, s);
}

lrt_status lrt_context_synchronize(lrt_context* ctx) {
    if (ctx == nullptr) return LRT_INVALID_ARGUMENT;
    const hipError_t e = hipStreamSynchronize(ctx->raster.stream());
    return finish(ctx, e == hipSuccess ? lrt::Status{}
                                       : lrt::Status::error(lrt::ErrorCode::DeviceFailure, hipGetErrorString(e)));
}

lrt_status lrt_copy_to_host(lrt_context* ctx, const void* device, void* host, size_t bytes) {
    if (ctx == nullptr) return LRT_INVALID_ARGUMENT;
    if (bytes == 0) return finish(ctx, {});
    if (device == nullptr || host == nullptr) {
        return finish(ctx, lrt::Status::error(lrt::ErrorCode::InvalidArgument, "null pointer"));
    }
    hipError_t e = hipMemcpyAsync(host, device, bytes, hipMemcpyDeviceToHost, ctx->raster.stream());
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->raster.stream());
    return finish(ctx, e == hipSuccess ? lrt::Status{}
                                       : lrt::Status::error(lrt::ErrorCode::DeviceFailure, hipGetErrorString(e)));
}


lrt_status lrt_write_exr(const char* path, uint32_t width, uint32_t height, const float* rgba, const float* depth,
                         int32_t half) {
    if (path == nullptr) return noContext(lrt::Status::error(lrt::ErrorCode::InvalidArgument, "null path"));
    return noContext(lrt::io::writeExr(path, width, height, rgba, depth, half != 0));
}

lrt_status lrt_splats_read(const char* path, lrt_raw_splats** out) {
    if (out == nullptr) return noContext(lrt::Status::error(lrt::ErrorCode::InvalidArgument, "null output"));
    *out = nullptr;
    if (path == nullptr) return noContext(lrt::Status::error(lrt::ErrorCode::InvalidArgument, "null path"));
    auto* raw = new (std::nothrow) lrt_raw_splats;
    if (raw == nullptr) return noContext(lrt::Status::error(lrt::ErrorCode::OutOfMemory, "raw splats"));
    const lrt::Status s = lrt::io::readSplats(path, raw->raw);
    if (!s.ok()) {
        delete raw;
        return noContext(s);
    }
    *out = raw;
    return noContext({});
}

lrt_status lrt_raw_splats_create(const float* records, uint32_t count, const lrt_splat_encoding* encoding,
                                 lrt_raw_splats** out) {
    using lrt::ErrorCode;
    using lrt::Status;
    if (out == nullptr) return noContext(Status::error(ErrorCode::InvalidArgument, "null output"));
    *out = nullptr;
    if (encoding == nullptr || (count > 0 && records == nullptr)) {
        return noContext(Status::error(ErrorCode::InvalidArgument, "null records or encoding"));
    }
    auto* raw = new (std::nothrow) lrt_raw_splats;
    if (raw == nullptr) return noContext(Status::error(ErrorCode::OutOfMemory, "raw splats"));
    const Status s = fromEncodingC(*encoding, raw->raw.encoding);
    if (!s.ok()) {
        delete raw;
        return noContext(s);
    }
    raw->raw.source = "records";
    raw->raw.count = count;
    raw->raw.records.assign(records, records + uint64_t{count} * encoding->floats_per_record);
    *out = raw;
    return noContext({});
}

void lrt_raw_splats_destroy(lrt_raw_splats* raw) { delete raw; }

uint32_t lrt_raw_splats_count(const lrt_raw_splats* raw) { return raw != nullptr ? raw->raw.count : 0u; }

lrt_status lrt_raw_splats_encoding(const lrt_raw_splats* raw, lrt_splat_encoding* out) {
    if (raw == nullptr || out == nullptr) return LRT_INVALID_ARGUMENT;
    *out = toEncodingC(raw->raw.encoding);
    return LRT_OK;
}

const float* lrt_raw_splats_records(const lrt_raw_splats* raw) {
    return raw != nullptr ? raw->raw.records.data() : nullptr;
}

uint32_t lrt_cloud_sh_words(const lrt_raw_splats* raw, uint32_t max_degree) {
    if (raw == nullptr) return 1u;
    const uint32_t keep = lrt::TileRasterizer::keptPerColour(raw->raw.encoding, max_degree);
    return keep == 0 ? 1u : (keep * 3 + 1) / 2;
}

lrt_status lrt_cloud_decode(lrt_context* ctx, const lrt_raw_splats* raw, uint32_t max_degree, float* positions,
                            uint32_t* shape, uint32_t* sh, lrt_cloud_info* info) {
    if (ctx == nullptr) return LRT_INVALID_ARGUMENT;
    if (raw == nullptr) return finish(ctx, lrt::Status::error(lrt::ErrorCode::InvalidArgument, "null raw splats"));
    lrt::TileRasterizer::CloudInfo out;
    const lrt::Status s = ctx->raster.decodeCloud(raw->raw, max_degree, positions, shape, sh, &out);
    if (s.ok() && info != nullptr) {
        info->declared = out.declared;
        info->count = out.count;
        info->rest_per_colour = out.restPerColour;
        info->sh_words = out.shWords;
        for (int k = 0; k < 3; ++k) {
            info->bounds_min[k] = out.boundsMin[k];
            info->bounds_max[k] = out.boundsMax[k];
        }
        info->upload_ms = out.uploadMs;
        info->decode_ms = out.decodeMs;
    }
    return finish(ctx, s);
}

void lrt_lod_build_settings_default(lrt_lod_build_settings* out) {
    if (out == nullptr) return;
    const lrt::lod::LodBuildSettings s;
    std::memset(out, 0, sizeof *out);
    out->max_group_fraction = s.maxGroupFraction;
    out->coarsest_level = s.coarsestLevel;
}

lrt_status lrt_lod_build(lrt_context* ctx, const lrt_splats* cloud, const lrt_lod_build_settings* settings,
                         lrt_lod_cloud** out) {
    if (ctx == nullptr) return LRT_INVALID_ARGUMENT;
    if (out == nullptr || cloud == nullptr) {
        return finish(ctx, lrt::Status::error(lrt::ErrorCode::InvalidArgument, "null argument"));
    }
    *out = nullptr;
    lrt::lod::LodBuildSettings s;
    if (settings != nullptr) {
        s.maxGroupFraction = settings->max_group_fraction;
        s.coarsestLevel = settings->coarsest_level;
    }
    auto* made = new (std::nothrow) lrt_lod_cloud();
    if (made == nullptr) return finish(ctx, lrt::Status::error(lrt::ErrorCode::OutOfMemory, "lrt_lod_cloud"));
    const lrt::Status st = lrt::lod::build(ctx->raster.device(), ctx->raster.stream(), toCloud(*cloud), s, made->lod);
    if (!st.ok()) {
        delete made;
        return finish(ctx, st);
    }
    *out = made;
    return finish(ctx, st);
}

void lrt_lod_destroy(lrt_lod_cloud* lod) { delete lod; }

lrt_status lrt_lod_get_info(const lrt_lod_cloud* lod, lrt_lod_info* out) {
    if (lod == nullptr || out == nullptr) return LRT_INVALID_ARGUMENT;
    std::memset(out, 0, sizeof *out);
    const lrt::lod::LodCloud& l = lod->lod;
    out->count = l.count;
    out->rest_per_colour = l.restPerColour;
    out->sh_words = l.shWords;
    out->levels = static_cast<uint32_t>(l.levels.size());
    out->merged = l.merged;
    for (size_t k = 0; k < l.levels.size() && k < 10; ++k) {
        out->level[k] = l.levels[k].level;
        out->first[k] = l.levels[k].first;
        out->groups[k] = l.levels[k].count;
    }
    for (int k = 0; k < 3; ++k) out->bounds_lo[k] = l.boundsLo[k];
    out->extent = l.extent;
    return LRT_OK;
}

namespace {
lrt_splats toSplatsC(const lrt::SplatCloud& c) {
    lrt_splats out{};
    out.positions = reinterpret_cast<const float*>(c.positions);
    out.shape = reinterpret_cast<const uint32_t*>(c.shape);
    out.sh = c.sh;
    out.filter3d = nullptr;
    out.count = c.count;
    out.rest_per_colour = c.restPerColour;
    out.sh_words = c.shWords;
    return out;
}
}   // namespace

lrt_status lrt_lod_splats(const lrt_lod_cloud* lod, lrt_splats* out, const uint32_t** groups) {
    if (lod == nullptr || out == nullptr) return LRT_INVALID_ARGUMENT;
    *out = toSplatsC(lod->lod.splats());
    if (groups != nullptr) *groups = lod->lod.groups.as<uint32_t>();
    return LRT_OK;
}

lrt_status lrt_lod_gaussians(const lrt_lod_cloud* lod, uint32_t level_index, lrt_splats* out,
                             const uint32_t** cells) {
    if (lod == nullptr || out == nullptr || level_index >= lod->lod.levels.size()) return LRT_INVALID_ARGUMENT;
    *out = toSplatsC(lod->lod.gaussians(level_index));
    if (cells != nullptr) *cells = lod->lod.cells.as<uint32_t>() + lod->lod.levels[level_index].first;
    return LRT_OK;
}

lrt_status lrt_lod_select(lrt_context* ctx, const lrt_projection* projection, const lrt_lod_instance* instances,
                          uint32_t count, float threshold, lrt_splat_instance* out, uint32_t* out_count,
                          lrt_cut_stats* stats) {
    if (ctx == nullptr) return LRT_INVALID_ARGUMENT;
    if (projection == nullptr || out_count == nullptr || (count > 0 && (instances == nullptr || out == nullptr))) {
        return finish(ctx, lrt::Status::error(lrt::ErrorCode::InvalidArgument, "null argument"));
    }
    *out_count = 0;
    std::vector<lrt::lod::LodInstance> list(count);
    for (uint32_t k = 0; k < count; ++k) {
        list[k].cloud = instances[k].cloud != nullptr ? &instances[k].cloud->lod : nullptr;
        list[k].objectToWorld = toMat(instances[k].object_to_world);
        list[k].edit = toEdit(instances[k].edit);
        if (instances[k].has_threshold != 0) list[k].threshold = instances[k].threshold;
    }
    std::vector<lrt::SplatInstance> drawn;
    std::vector<lrt::lod::CutStats> cs;
    const lrt::Status s = ctx->cut.select(ctx->raster.device(), ctx->raster.stream(), toProjection(*projection), list,
                                          threshold, drawn, &cs);
    if (!s.ok()) return finish(ctx, s);
    ctx->cutClouds.resize(drawn.size());
    for (size_t k = 0; k < drawn.size(); ++k) {
        ctx->cutClouds[k] = toSplatsC(*drawn[k].splats);
        out[k] = lrt_splat_instance{};
        out[k].splats = &ctx->cutClouds[k];
        for (int e = 0; e < 16; ++e) out[k].object_to_world[e] = drawn[k].objectToWorld.m[static_cast<size_t>(e)];
    }
    // a drawn instance keeps its own edit: the k-th drawn is the k-th instance that drew something
    size_t next = 0;
    for (uint32_t k = 0; k < count && next < drawn.size(); ++k) {
        if (k < cs.size() && cs[k].splats + cs[k].merged > 0) out[next++].edit = instances[k].edit;
    }
    if (stats != nullptr) {
        for (uint32_t k = 0; k < count; ++k) {
            stats[k] = lrt_cut_stats{};
            if (k < cs.size()) stats[k] = lrt_cut_stats{cs[k].splats, cs[k].merged, cs[k].available};
        }
    }
    *out_count = static_cast<uint32_t>(drawn.size());
    return finish(ctx, s);
}

}   // extern "C"
