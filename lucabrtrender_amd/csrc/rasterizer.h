// The C++20 host side of the splat tile rasteriser: the shape of the
// reference's lrt::render::TileRasterizer (modules/render/include/lrt/render/
// TileRasterizer.h:136-193) over HIP instead of slang-rhi. It owns every
// scratch buffer (grown by half again from 1024 and never shrunk, as the
// reference's reserve* do, TileRasterizer.cpp:26-32,105-177), queues the
// kernels of kernels.h on one stream and waits where the reference waits.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <span>
#include <string>

#include "frame.h"
#include "kernels.h"
#include "readers.h"
#include "status.h"

namespace lrt {

// Host math in FP64 (aofx Transform.h): row-major 4x4 acting on column vectors.
struct Mat4 {
    std::array<double, 16> m{1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    [[nodiscard]] double at(int r, int c) const { return m[static_cast<size_t>(r * 4 + c)]; }
    double& at(int r, int c) { return m[static_cast<size_t>(r * 4 + c)]; }
};
struct Vec3 {
    double x = 0, y = 0, z = 0;
};

Mat4 operator*(const Mat4& a, const Mat4& b);
Mat4 inverseAffine(const Mat4& m);
Mat4 scaling(const Vec3& s);
Mat4 rotationZ(double degrees);
Mat4 lookAt(const Vec3& eye, const Vec3& target, const Vec3& up);
Vec3 transformPoint(const Mat4& m, const Vec3& p);

struct Lens {   // modules/render/include/lrt/render/Camera.h:23-44
    bool orthographic = false;
    double focal = 50.0, haperture = 24.576, vaperture = 18.672, nearZ = 0.1, farZ = 10000.0;
    double windowTranslate[2] = {0.0, 0.0};
    double windowScale[2] = {1.0, 1.0};
    double windowRoll = 0.0;
};

struct Camera {   // Camera.h:46-53
    Mat4 cameraToWorld;
    Lens lens;
    static Camera lookingAt(const Vec3& eye, const Vec3& target, const Vec3& up = {0.0, 1.0, 0.0});
};

struct Projection {   // Camera.h:56-80
    Mat4 worldToView;
    double focalX = 0, focalY = 0, centreX = 0, centreY = 0, nearZ = 0.1, farZ = 10000.0;
    bool orthographic = false;
    Vec3 eyeWorld;
};

Projection projectionFor(const Camera& camera, uint32_t width, uint32_t height);
uint32_t bitsFor(uint32_t values);

struct SplatCloud {   // scene::GpuSplats, device pointers the caller owns
    const float4* positions = nullptr;
    const uint4* shape = nullptr;
    const uint32_t* sh = nullptr;
    const float* filter3d = nullptr;
    uint32_t count = 0, restPerColour = 0, shWords = 1;
};

struct SplatInstance {   // render::SplatInstance without relighting
    const SplatCloud* splats = nullptr;
    Mat4 objectToWorld;
    SplatEditParams edit;
};

struct PointCloud {   // scene::GpuPoints (GpuClouds.h:78-85), device pointers the caller owns
    const float4* positions = nullptr;   // x, y, z, 1
    const uint32_t* colours = nullptr;   // two words a point: f16 r | g, b | a (linear); 8-byte aligned
    uint32_t count = 0;
};

struct PointStyle {   // render::PointStyle (Points.h:21-33), the fields the disc route reads
    enum class Size { World, Pixels };
    Size sizeMode = Size::World;
    float size = 0.01F;                 // diameter, world units or pixels
    bool constantColour = false;
    std::array<float, 3> colour{1.0F, 1.0F, 1.0F};
};

struct PointInstance {   // render::PointInstance (Points.h:35-39)
    const PointCloud* points = nullptr;
    Mat4 objectToWorld;
    PointStyle style;
};

struct RenderSettings {   // render::RenderSettings (TileRasterizer.h:105-118)
    uint32_t width = 1920, height = 1080;
    bool antialias = true;
    uint32_t maxShDegree = 3;
    std::array<float, 4> background{0, 0, 0, 0};
    uint32_t depthMode = 0;
    float depthThreshold = 0.5F;
    bool linearise = true;
    bool timeStages = false;
    float dilation = 0.3F;
    bool filter3d = false;
    uint32_t blendBatch = 256;          // records each quadrant's LDS list holds: 64, 128 or 256
    uint32_t blendSegment = 0;          // pairs a quadrant's wave walks before its tile is split (0: never)
    bool keepWalkState = false;         // keep each pixel's last contributor and final transmittance (debug)
    bool heaviestTilesFirst = true;     // launch the tiles with the most pairs first
    bool persistentBlend = false;       // batches: a persistent blend taking quadrants from a queue
    bool useGraph = false;              // batches: replay the launch sequence as a captured hipGraph
    uint32_t tileSize = 16;             // pixels a square tile's side: 8, 16 (the reference's) or 32
    uint32_t radixBits = 0;             // the sort's widest digit: 8, 9 or 10 (0: 9 where it saves a pass, else 8)
};

struct RenderTargets {   // caller-owned device images, bottom row first
    uint32_t width = 0, height = 0;
    float4* colour = nullptr;
    float* depth = nullptr;
};

struct FrameStats {   // render::FrameStats (TileRasterizer.h:120-126)
    uint32_t splats = 0, visible = 0, pairs = 0;
    // The sort's key: tile bits plus the bits of the frame's depth-key range, and its passes.
    uint32_t keyBits = 0, sortPasses = 0;
    double projectMs = 0, depthSortMs = 0, countsMs = 0, emitMs = 0, tileSortMs = 0, blendMs = 0, totalMs = 0;
};

// One camera of a batch: its projection and its images (the targets' size is
// the view's frame size).
struct View {
    Projection projection;
    RenderTargets targets;
};

// setFrame + setObject (modules/render/src/FrameParams.h:32-69).
FrameParams frameParamsFor(const Projection& projection, const SplatInstance& instance,
                           const RenderSettings& settings, uint32_t base);

// A device buffer that grows by half again from 1024 elements and never shrinks.
class DeviceBuffer {
public:
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    ~DeviceBuffer();
    Status reserve(uint64_t bytes, const char* label);
    [[nodiscard]] void* get() const noexcept { return ptr_; }
    template <typename T>
    [[nodiscard]] T* as() const noexcept {
        return static_cast<T*>(ptr_);
    }
    [[nodiscard]] uint64_t bytes() const noexcept { return bytes_; }

private:
    void* ptr_ = nullptr;
    uint64_t bytes_ = 0;
};

class TileRasterizer {
public:
    TileRasterizer() = default;
    TileRasterizer(const TileRasterizer&) = delete;
    TileRasterizer& operator=(const TileRasterizer&) = delete;
    ~TileRasterizer();

    Status create(int device, hipStream_t stream);
    void setStream(hipStream_t stream) { stream_ = stream; }
    [[nodiscard]] hipStream_t stream() const noexcept { return stream_; }
    [[nodiscard]] int device() const noexcept { return device_; }

    // TileRasterizer::render (TileRasterizer.cpp:188-484). Point instances are
    // drawn as opaque discs sorted with the splats (their slots after the
    // splats'); `under` (the frame's size, colour and depth both set) is what
    // lies under the splats and ends each pixel's walk at its depth. Either
    // one selects the composite blend.
    Status render(const Projection& projection, std::span<const SplatInstance> instances,
                  const RenderSettings& settings, RenderTargets& targets, FrameStats* stats,
                  std::span<const PointInstance> points = {}, const RenderTargets* under = nullptr);
    Status render(const Camera& camera, std::span<const SplatInstance> instances, const RenderSettings& settings,
                  RenderTargets& targets, FrameStats* stats, std::span<const PointInstance> points = {},
                  const RenderTargets* under = nullptr) {
        return render(projectionFor(camera, settings.width, settings.height), instances, settings, targets, stats,
                      points, under);
    }

    // Many views of the same instances in one launch sequence: one projection
    // launch per instance for all views, one scan, one duplication, one sort
    // over every view's pairs, one blend over every view's tiles. No host wait
    // until the end (the pair count stays on the device); with
    // settings.useGraph the sequence is captured once per batch shape and
    // replayed. settings.width/height are ignored (each view's targets give
    // its size). stats: the batch's totals.
    Status renderViews(std::span<const View> views, std::span<const SplatInstance> instances,
                       const RenderSettings& settings, FrameStats* stats);
    // Times the batch under each launch choice (list length 128/256, grid or
    // persistent blend, graph replay or not; the best of three runs each) and
    // leaves the fastest in `settings`; *ms gets its time.
    Status autotuneViews(std::span<const View> views, std::span<const SplatInstance> instances,
                         RenderSettings& settings, float* ms);

    // Many views of one FP32-parameter cloud: renderViews' launch sequence with
    // renderDiff's differentiable stages (pairs carry their slot-major index,
    // each pixel keeps its raw rgb, last contributor and final transmittance),
    // for backwardViews(). Each view's images are those renderDiff would draw
    // for it alone, bit for bit.
    Status renderDiffViews(std::span<const View> views, const kern::DiffCloud& cloud, const Mat4& objectToWorld,
                           const RenderSettings& settings, FrameStats* stats);
    // Gradients of the last renderDiffViews' parameters, summed over its views
    // in view order, from each view's dL/d(colour) (device, one float4 a pixel
    // of that view, bottom row first). millis as backward()'s.
    Status backwardViews(std::span<const float4* const> gradColours, const kern::DiffGrads& grads, float* millis);

    // The differentiable frame: one FP32-parameter cloud, drawn by the same
    // stages as render(); it keeps what backward() needs.
    Status renderDiff(const Projection& projection, const kern::DiffCloud& cloud, const Mat4& objectToWorld,
                      const RenderSettings& settings, RenderTargets& targets, FrameStats* stats);
    // The brute force in FP64 (the error oracle of whole frames): every splat of
    // every instance in (view z, index) order, or with keyed in the
    // rasteriser's (24-bit depth key, slot) order, per pixel, no cap. colour: 4
    // doubles a pixel, depth 1 (may be null), device, settings' size.
    // With pixels (device, pixelCount entries y * width + x) only those pixels,
    // output k for pixels[k].
    Status referenceRender(const Projection& projection, std::span<const SplatInstance> instances,
                           const RenderSettings& settings, double* colour, double* depth, bool keyed,
                           const uint32_t* pixels = nullptr, uint32_t pixelCount = 0);

    // Gradients of the last renderDiff's parameters given dL/d(colour) (one
    // float4 a pixel, bottom row first). millis (may be null) receives the GPU
    // ms of the blend backward and of the preprocess backward.
    Status backward(const float4* gradColour, const kern::DiffGrads& grads, float* millis);

    // PrefixSum::apply and RadixSort::sort over caller buffers, on this stream
    // (the sort's digits at most radixBits wide: 8, the reference's, to 10).
    Status prefixSum(const uint32_t* in, uint32_t* out, uint32_t* total, uint32_t count);
    Status radixSort(uint64_t* keys, uint32_t* values, uint64_t* scratchKeys, uint32_t* scratchValues,
                     uint32_t count, uint32_t keyBits, bool* resultInScratch, uint32_t radixBits = 8);

    // What the last render left (LRT_DEBUG_* of lrt_hip.h).
    Status debugCopy(int which, void* host, size_t bytes, size_t* needed);

    // CloudLoader::upload (modules/scene/src/GpuClouds.cpp:199-227): raw
    // records decoded into packed splats at caller device buffers sized for
    // raw.count (positions 16 B, shape 16 B, sh shWords u32 a splat; sh may be
    // null when no rest coefficient is kept), in slices of at most 256 MiB of
    // records, compacted in file order. Synchronous. Reuses the frame's
    // per-splat scratch, so the last frame's debug buffers are gone after it.
    struct CloudInfo {
        uint32_t declared = 0, count = 0, restPerColour = 0, shWords = 1;
        float boundsMin[3] = {0, 0, 0}, boundsMax[3] = {0, 0, 0};
        float uploadMs = 0, decodeMs = 0;   // GPU time: the records' copies in, and the kernels
    };
    static uint32_t keptPerColour(const io::SplatEncoding& e, uint32_t maxDegree);
    Status decodeCloud(const io::RawSplats& raw, uint32_t maxDegree, float* positions, uint32_t* shape, uint32_t* sh,
                       CloudInfo* info);

private:
    Status reserveSplats(uint32_t count);
    Status reservePairs(uint32_t count, uint32_t keyBits, bool training = false, uint32_t radixBits = 0);
    Status reserveTargets(uint32_t width, uint32_t height, uint32_t tiles, bool training);
    Status checkFrame(const RenderSettings& settings, const RenderTargets& targets) const;
    // A batch of views of the instances, or (diff non-null, one instance giving
    // the cloud's place) of an FP32-parameter cloud kept for backwardViews().
    Status runViews(std::span<const View> views, std::span<const SplatInstance> instances,
                    const kern::DiffCloud* diff, const RenderSettings& settings, FrameStats* stats);
    // The launch sequence of a batch (no waits), on `stream`.
    void enqueueViews(std::span<const SplatInstance> instances, const kern::DiffCloud* diff, uint32_t viewCount,
                      uint32_t splatsPerView, uint32_t tiles, const RenderSettings& settings, hipStream_t stream);
    // Everything after the projection: scan, duplicate, sort, ranges, blend
    // (the composite blend when `composite`, over `under` when non-null).
    // scanZeroed: the projection already cleared the scan's look-back words
    // and the counters.
    Status finishFrame(const FrameParams& first, uint32_t splatCount, const RenderSettings& settings,
                       RenderTargets& targets, FrameStats* stats, bool training, bool composite = false,
                       const RenderTargets* under = nullptr, bool scanZeroed = false);
    Status check(hipError_t e, const char* what);

    int device_ = 0;
    uint32_t computeUnits_ = 1;
    hipStream_t stream_ = nullptr;
    bool created_ = false;

    // Per slot.
    uint32_t splatCapacity_ = 0;
    DeviceBuffer proj_, rects_, touched_, offsets_, scanStatus_;
    // The projecting workgroups' depth-key ranges (DepthBlocks), two words each.
    DeviceBuffer depthBlocks_;
    uint32_t depthEntries_ = 0;   // the entries the frame's projections wrote
    kern::TileDepthLayout viewsLayout_;   // the batch's sort key depth field (from its near and far planes)
    // Per pair, ping-pong.
    uint32_t pairCapacity_ = 0;
    DeviceBuffer keys_[2], values_[2], sortWork_;
    // Training frames: each pair's slot in slot-major order, the slots in sorted
    // order, and the backward's per-pair gradient records.
    DeviceBuffer pairSlot_, sortedSlots_, pairGrad_;
    // The FP64 brute force's records, keys and values (ping-pong) and sort workspace.
    DeviceBuffer refRecords_, refKeys_[2], refValues_[2], refWork_;
    DeviceBuffer loadRaw_, loadBox_;   // a slice of raw records; the bounds' partials and box
    uint64_t sortWorkWords_ = 0;
    // Per tile and per pixel.
    uint32_t tileCapacity_ = 0;
    uint64_t pixelCapacity_ = 0;
    DeviceBuffer ranges_, tileOrder_, lastCount_, finalT_, accum_;
    // Split blends: the plan's tasks and split tiles, and the segments' partial walks.
    DeviceBuffer planTasks_, planCombine_, partials_;
    // Counters: [0] pairs, [1] visible, [2] overflow, [3] unused, [4] sort error, [5] scan error,
    // [6..7] the split blend's plan counts (the views' queue), [8] least and [9] greatest depth key
    // of the frame's drawn slots, [10] drawn slots outside that range.
    DeviceBuffer counters_;
    uint32_t* hostCounters_ = nullptr;   // pinned
    hipEvent_t events_[9] = {};
    hipEvent_t countsCopied_ = nullptr;   // the pair-count readback has landed (no timing)
    // GPU milliseconds of the last timed render: project, scan, duplicate, sort
    // histogram, sort passes, ranges (with the heaviest-first ordering), blend, their sum.
    float kernelMs_[8] = {};

    // What the last render left.
    uint32_t lastSplats_ = 0, lastVisible_ = 0, lastPairs_ = 0, lastTiles_ = 0;
    bool lastOrdered_ = false;
    bool lastWalkDropped_ = false;   // the last frame kept no per-pixel walk state
    uint64_t lastPixels_ = 0;
    int lastSorted_ = 0;   // which of keys_/values_ hold the sorted pairs
    bool lastNarrow_ = false;       // the sorted keys are u32 (kern::sortNarrows)
    uint32_t lastTileShift_ = 0;    // their tile is key >> lastTileShift_
    bool lastSlotsSorted_ = false;  // the sorted values are pair indices (their slots: pairSlot_)
    // The last plain (one view, splats only, not split) blend's launch, for
    // the walk-step count of LRT_DEBUG_WALK_STEPS.
    bool lastBlendPlain_ = false;
    FrameParams lastBlendFrame_{};
    uint32_t lastBlendList_ = 0, lastBlendSlots_ = 0;
    const uint32_t* lastBlendOrder_ = nullptr;
    const uint32_t* lastBlendPairSlots_ = nullptr;
    DeviceBuffer walkSteps_;
    // Batches: the frame constants and view records (pinned host copies that
    // a launch uploads, and their device arrays), the capture stream and the
    // graph of the last batch shape.
    FrameParams* hostParams_ = nullptr;
    ViewInfo* hostViews_ = nullptr;
    uint32_t batchCapacity_ = 0;   // FrameParams entries the four buffers hold
    DeviceBuffer viewParams_, viewInfo_;
    hipStream_t captureStream_ = nullptr;
    hipGraphExec_t graph_ = nullptr;
    std::string graphKey_;
    // The last differentiable frame, or batch of views (its view count).
    bool diffReady_ = false;
    bool diffViewsReady_ = false;
    uint32_t diffViews_ = 0;
    kern::DiffCloud diffCloud_{};
    FrameParams diffFrame_{};
    // A batch backward's colour-gradient pointers, one a view (pinned host, device).
    const float4** hostGradPtrs_ = nullptr;
    uint32_t gradPtrCapacity_ = 0;
    DeviceBuffer viewGrads_;
};

}   // namespace lrt
