// TileRasterizer::render over HIP: the reference's frame
// (modules/render/src/TileRasterizer.cpp:188-484) with its two sorts replaced
// by one stable sort of 64-bit (tile, depth) keys.
//
//   project (per instance) -> scan of tiles touched (pairs + visible count)
//   -> [one wait: the pair count sizes the pair buffers, as the reference's
//      readback at :385-397 does] -> duplicate -> sort -> ranges -> blend
//   -> [the final wait, as :476].
#include <hip/hip_runtime.h>

#include <algorithm>
#include <bit>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <limits>
#include <string>

#include "kernels.h"
#include "rasterizer.h"

namespace lrt {
namespace {

uint64_t grow(uint64_t have, uint64_t need) {
    uint64_t size = std::max<uint64_t>(have, 1024);
    while (size < need) size = size * 3 / 2;
    return size;
}

// The sort keeps 30-bit counts in its look-back words.
constexpr uint32_t kMaxPairs = (1u << 30) - 1;

}   // namespace

DeviceBuffer::~DeviceBuffer() {
    if (ptr_ != nullptr) (void)hipFree(ptr_);
}

Status DeviceBuffer::reserve(uint64_t bytes, const char* label) {
    if (bytes <= bytes_) return {};
    if (ptr_ != nullptr) {
        (void)hipFree(ptr_);
        ptr_ = nullptr;
        bytes_ = 0;
    }
    const hipError_t e = hipMalloc(&ptr_, std::max<uint64_t>(bytes, 256));
    if (e != hipSuccess) {
        ptr_ = nullptr;
        return Status::error(e == hipErrorOutOfMemory ? ErrorCode::OutOfMemory : ErrorCode::DeviceFailure,
                             std::string("allocating ") + label + " (" + std::to_string(bytes) +
                                 " bytes): " + hipGetErrorString(e));
    }
    bytes_ = bytes;
    return {};
}

TileRasterizer::~TileRasterizer() {
    if (!created_) return;
    (void)hipSetDevice(device_);
    for (hipEvent_t& e : events_)
        if (e != nullptr) (void)hipEventDestroy(e);
    if (countsCopied_ != nullptr) (void)hipEventDestroy(countsCopied_);
    if (hostCounters_ != nullptr) (void)hipHostFree(hostCounters_);
    if (hostParams_ != nullptr) (void)hipHostFree(hostParams_);
    if (hostViews_ != nullptr) (void)hipHostFree(hostViews_);
    if (hostGradPtrs_ != nullptr) (void)hipHostFree(hostGradPtrs_);
    if (graph_ != nullptr) (void)hipGraphExecDestroy(graph_);
    if (captureStream_ != nullptr) (void)hipStreamDestroy(captureStream_);
}

Status TileRasterizer::check(hipError_t e, const char* what) {
    if (e == hipSuccess) return {};
    return Status::error(e == hipErrorOutOfMemory ? ErrorCode::OutOfMemory : ErrorCode::DeviceFailure,
                         std::string(what) + ": " + hipGetErrorString(e));
}

Status TileRasterizer::create(int device, hipStream_t stream) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        return Status::error(ErrorCode::DeviceFailure, "no HIP device");
    }
    if (device < 0 || device >= count) {
        return Status::error(ErrorCode::InvalidArgument, "device " + std::to_string(device) + " of " +
                                                             std::to_string(count));
    }
    device_ = device;
    stream_ = stream;
    Status s = check(hipSetDevice(device_), "hipSetDevice");
    if (!s.ok()) return s;
    for (hipEvent_t& e : events_) {
        s = check(hipEventCreate(&e), "hipEventCreate");
        if (!s.ok()) return s;
    }
    s = check(hipEventCreateWithFlags(&countsCopied_, hipEventDisableTiming), "hipEventCreate");
    if (!s.ok()) return s;
    int cus = 0;
    s = check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device_), "device attribute");
    if (!s.ok()) return s;
    computeUnits_ = static_cast<uint32_t>(cus > 0 ? cus : 1);
    s = check(hipHostMalloc(reinterpret_cast<void**>(&hostCounters_), 64, hipHostMallocDefault), "hipHostMalloc");
    if (!s.ok()) return s;
    s = counters_.reserve(64, "counters");
    if (!s.ok()) return s;
    kern::fill32(counters_.as<uint32_t>(), 0u, 16, stream_);
    created_ = true;
    return check(hipStreamSynchronize(stream_), "create");
}

Status TileRasterizer::reserveSplats(uint32_t count) {
    if (count <= splatCapacity_) return {};
    const uint64_t n = std::min<uint64_t>(grow(splatCapacity_, count), 0xFFFFFFFFull);
    Status s;
    if (!(s = proj_.reserve(n * sizeof(ProjRecord), "splat.proj")).ok()) return s;
    if (!(s = rects_.reserve(n * 8, "splat.rects")).ok()) return s;
    if (!(s = touched_.reserve(n * 4, "splat.touched")).ok()) return s;
    if (!(s = offsets_.reserve(n * 4, "splat.offsets")).ok()) return s;
    if (!(s = scanStatus_.reserve(kern::scanStatusWords(static_cast<uint32_t>(n)) * 8, "splat.scanStatus"))
             .ok())
        return s;
    splatCapacity_ = static_cast<uint32_t>(n);
    return {};
}

Status TileRasterizer::reservePairs(uint32_t count, uint32_t keyBits, bool training, uint32_t radixBits) {
    Status s;
    if (count > pairCapacity_) {
        const uint64_t n = std::min<uint64_t>(grow(pairCapacity_, count), kMaxPairs);
        for (int k = 0; k < 2; ++k) {
            if (!(s = keys_[k].reserve(n * 8, "pair.keys")).ok()) return s;
            if (!(s = values_[k].reserve(n * 4, "pair.values")).ok()) return s;
        }
        pairCapacity_ = static_cast<uint32_t>(n);
    }
    if (training) {
        if (!(s = pairSlot_.reserve(uint64_t{pairCapacity_} * 4, "pair.slot")).ok()) return s;
        if (!(s = sortedSlots_.reserve(uint64_t{pairCapacity_} * 4, "pair.sortedSlots")).ok()) return s;
    }
    const uint64_t words = kern::sortWorkspaceWordsUpTo(pairCapacity_, std::max<uint32_t>(keyBits, 40), radixBits);
    if (words > sortWorkWords_) {
        if (!(s = sortWork_.reserve(words * 4, "pair.sortWorkspace")).ok()) return s;
        sortWorkWords_ = words;
    }
    return {};
}

Status TileRasterizer::reserveTargets(uint32_t width, uint32_t height, uint32_t tiles, bool training) {
    Status s;
    if (tiles > tileCapacity_) {
        if (!(s = ranges_.reserve(uint64_t{tiles} * 8, "tile.ranges")).ok()) return s;
        if (!(s = tileOrder_.reserve(uint64_t{tiles} * 4, "tile.order")).ok()) return s;
        tileCapacity_ = tiles;
    }
    const uint64_t pixels = uint64_t{width} * height;
    if (pixels > pixelCapacity_) {
        if (!(s = lastCount_.reserve(pixels * 4, "pixel.lastCount")).ok()) return s;
        if (!(s = finalT_.reserve(pixels * 4, "pixel.finalT")).ok()) return s;
        pixelCapacity_ = pixels;
    }
    if (training && !(s = accum_.reserve(pixels * 16, "pixel.accum")).ok()) return s;
    return {};
}

Status TileRasterizer::checkFrame(const RenderSettings& settings, const RenderTargets& targets) const {
    if (!created_) return Status::error(ErrorCode::InvalidArgument, "the rasteriser was not created");
    if (settings.width == 0 || settings.height == 0) {
        return Status::error(ErrorCode::InvalidArgument, "a frame needs a size");
    }
    if (settings.tileSize != 8 && settings.tileSize != 16 && settings.tileSize != 32) {
        return Status::error(ErrorCode::InvalidArgument, "a tile size other than 8, 16 or 32 pixels");
    }
    if (settings.radixBits != 0 && (settings.radixBits < 8 || settings.radixBits > 10)) {
        return Status::error(ErrorCode::InvalidArgument, "a sort digit other than 8, 9 or 10 bits (or 0)");
    }
    const uint32_t shift = tileShiftOf(settings.tileSize);
    if (tilesAlong(settings.width, shift) > 0xFFFFu || tilesAlong(settings.height, shift) > 0xFFFFu) {
        return Status::error(ErrorCode::Unsupported, "a frame over 65535 tiles wide or tall");
    }
    if (targets.width != settings.width || targets.height != settings.height || targets.colour == nullptr) {
        return Status::error(ErrorCode::InvalidArgument, "the targets are not the frame's size");
    }
    return {};
}

Status TileRasterizer::render(const Projection& projection, std::span<const SplatInstance> instances,
                              const RenderSettings& settings, RenderTargets& targets, FrameStats* statsOut,
                              std::span<const PointInstance> points, const RenderTargets* under) {
    Status s = checkFrame(settings, targets);
    if (!s.ok()) return s;
    if (under != nullptr && (under->width != settings.width || under->height != settings.height)) {
        return Status::error(ErrorCode::InvalidArgument, "the under-layer is not the frame's size");
    }
    if (under != nullptr && (under->colour == nullptr || under->depth == nullptr)) {
        return Status::error(ErrorCode::InvalidArgument, "an under-layer needs both colour and depth");
    }
    uint64_t total = 0;
    for (const SplatInstance& instance : instances) {
        if (instance.splats == nullptr) continue;
        const SplatCloud& c = *instance.splats;
        if (c.count > 0 && (c.positions == nullptr || c.shape == nullptr)) {
            return Status::error(ErrorCode::InvalidArgument, "a cloud without positions or shape");
        }
        if (c.count > 0 && c.restPerColour > 0 && c.sh == nullptr) {
            return Status::error(ErrorCode::InvalidArgument, "a cloud with harmonics but no sh buffer");
        }
        if (c.restPerColour != 0 && c.restPerColour != 3 && c.restPerColour != 8 && c.restPerColour != 15) {
            return Status::error(ErrorCode::InvalidArgument, "restPerColour must be 0, 3, 8 or 15");
        }
        if (c.restPerColour > 0 && c.shWords < (c.restPerColour * 3 + 1) / 2) {
            return Status::error(ErrorCode::InvalidArgument, "shWords too small for restPerColour");
        }
        if (c.restPerColour > 0 && c.shWords > 23) {
            return Status::error(ErrorCode::Unsupported, "more than 23 harmonic words a splat");
        }
        if (settings.filter3d && c.count > 0 && c.filter3d == nullptr) {
            return Status::error(ErrorCode::InvalidArgument, "the 3D filter is on and a cloud has no filter3d");
        }
        total += c.count;
    }
    for (const PointInstance& instance : points) {
        if (instance.points == nullptr) continue;
        const PointCloud& c = *instance.points;
        if (c.count > 0 && c.positions == nullptr) {
            return Status::error(ErrorCode::InvalidArgument, "a point cloud without positions");
        }
        if (c.count > 0 && !instance.style.constantColour && c.colours == nullptr) {
            return Status::error(ErrorCode::InvalidArgument, "a point cloud without colours and no constant colour");
        }
        if (!(instance.style.size >= 0.0F)) {
            return Status::error(ErrorCode::InvalidArgument, "a point size below 0");
        }
        total += c.count;
    }
    if (total > 0xFFFFFFFFull) return Status::error(ErrorCode::Unsupported, "more than 2^32 splats in one frame");
    const uint32_t splatCount = static_cast<uint32_t>(total);
    const uint32_t shift = tileShiftOf(settings.tileSize);
    const uint32_t tiles = tilesAlong(settings.width, shift) * tilesAlong(settings.height, shift);
    if (!(s = check(hipSetDevice(device_), "hipSetDevice")).ok()) return s;
    if (!(s = reserveSplats(std::max(splatCount, 1u))).ok()) return s;
    if (!(s = reserveTargets(settings.width, settings.height, tiles, false)).ok()) return s;

    // One depth-key range entry for each projecting workgroup of the frame.
    uint64_t entries = 0;
    for (const SplatInstance& instance : instances) {
        if (instance.splats != nullptr) entries += projectBlocks(instance.splats->count);
    }
    for (const PointInstance& instance : points) {
        if (instance.points != nullptr) entries += projectBlocks(instance.points->count);
    }
    if (!(s = depthBlocks_.reserve(std::max<uint64_t>(entries, 1) * 8, "frame.depthBlocks")).ok()) return s;
    DepthBlocks depth{depthBlocks_.as<uint32_t>(), 0};

    if (settings.timeStages) (void)hipEventRecord(events_[0], stream_);
    // Project every instance into its own slots.
    uint32_t base = 0;
    FrameParams first{};
    bool haveFirst = false;
    // The scan's look-back words and the counters are cleared by the first
    // projection launch rather than by a launch of their own.
    uint32_t* counters = counters_.as<uint32_t>();
    bool scanZeroed = false;
    for (const SplatInstance& instance : instances) {
        const SplatCloud* cloud = instance.splats;
        FrameParams f = frameParamsFor(projection, instance, settings, base);
        if (!haveFirst) {
            first = f;
            haveFirst = true;
        }
        if (cloud == nullptr || cloud->count == 0) continue;
        kern::ZeroSpans zero;
        if (!scanZeroed) {
            zero = kern::ZeroSpans{reinterpret_cast<uint32_t*>(scanStatus_.as<uint64_t>()),
                                   kern::scanZeroWords(splatCount), counters, 12};
            scanZeroed = true;
        }
        kern::project(f, cloud->positions, cloud->shape, cloud->sh, cloud->filter3d, proj_.as<float4>(),
                      rects_.as<uint2>(), touched_.as<uint32_t>(), depth, stream_, zero);
        depth.first += projectBlocks(cloud->count);
        base += cloud->count;
    }
    // Then the points, as discs, in the slots after the splats'.
    for (const PointInstance& instance : points) {
        const PointCloud* cloud = instance.points;
        if (cloud == nullptr || cloud->count == 0) continue;
        SplatInstance placed;
        placed.objectToWorld = instance.objectToWorld;
        FrameParams f = frameParamsFor(projection, placed, settings, base);
        f.count = cloud->count;
        PointParams style{};
        style.sizeMode = instance.style.sizeMode == PointStyle::Size::Pixels ? 1u : 0u;
        style.size = instance.style.size;
        style.constantColour = instance.style.constantColour ? 1u : 0u;
        for (int k = 0; k < 3; ++k) style.colour[k] = instance.style.colour[static_cast<size_t>(k)];
        kern::projectPoints(f, style, cloud->positions, cloud->colours, proj_.as<float4>(), rects_.as<uint2>(),
                            touched_.as<uint32_t>(), depth, stream_);
        depth.first += projectBlocks(cloud->count);
        base += cloud->count;
    }
    depthEntries_ = depth.first;
    if (!haveFirst) {
        SplatInstance none;
        first = frameParamsFor(projection, none, settings, 0);
    }
    diffReady_ = false;
    diffViewsReady_ = false;
    lastBlendPlain_ = false;
    const bool composite = under != nullptr || !points.empty();
    return finishFrame(first, splatCount, settings, targets, statsOut, false, composite, under, scanZeroed);
}

Status TileRasterizer::renderDiff(const Projection& projection, const kern::DiffCloud& cloud,
                                  const Mat4& objectToWorld, const RenderSettings& settings, RenderTargets& targets,
                                  FrameStats* statsOut) {
    Status s = checkFrame(settings, targets);
    if (!s.ok()) return s;
    if (settings.tileSize != kTile) {
        return Status::error(ErrorCode::Unsupported, "a differentiable frame with tiles other than 16 pixels");
    }
    if (cloud.count > 0 && (cloud.means == nullptr || cloud.logScales == nullptr || cloud.quats == nullptr ||
                            cloud.logits == nullptr || cloud.sh == nullptr)) {
        return Status::error(ErrorCode::InvalidArgument, "a differentiable cloud with a missing parameter");
    }
    if (cloud.restPerColour != 0 && cloud.restPerColour != 3 && cloud.restPerColour != 8 &&
        cloud.restPerColour != 15) {
        return Status::error(ErrorCode::InvalidArgument, "restPerColour must be 0, 3, 8 or 15");
    }
    if (settings.filter3d && cloud.count > 0 && cloud.filter3d == nullptr) {
        return Status::error(ErrorCode::InvalidArgument, "the 3D filter is on and the cloud has no filter3d");
    }
    const uint32_t shift = tileShiftOf(settings.tileSize);
    const uint32_t tiles = tilesAlong(settings.width, shift) * tilesAlong(settings.height, shift);
    if (!(s = check(hipSetDevice(device_), "hipSetDevice")).ok()) return s;
    if (!(s = reserveSplats(std::max(cloud.count, 1u))).ok()) return s;
    if (!(s = reserveTargets(settings.width, settings.height, tiles, true)).ok()) return s;

    SplatCloud shapeOnly;
    shapeOnly.count = cloud.count;
    shapeOnly.restPerColour = cloud.restPerColour;
    SplatInstance instance;
    instance.splats = &shapeOnly;
    instance.objectToWorld = objectToWorld;
    const FrameParams f = frameParamsFor(projection, instance, settings, 0);
    if (!(s = depthBlocks_.reserve(std::max<uint64_t>(projectBlocks(cloud.count), 1) * 8, "frame.depthBlocks")).ok())
        return s;
    if (settings.timeStages) (void)hipEventRecord(events_[0], stream_);
    kern::projectDiff(f, cloud, proj_.as<float4>(), rects_.as<uint2>(), touched_.as<uint32_t>(),
                      DepthBlocks{depthBlocks_.as<uint32_t>(), 0}, stream_);
    depthEntries_ = projectBlocks(cloud.count);
    diffReady_ = false;
    diffViewsReady_ = false;
    lastBlendPlain_ = false;
    if (!(s = finishFrame(f, cloud.count, settings, targets, statsOut, true)).ok()) return s;
    diffCloud_ = cloud;
    diffFrame_ = f;
    diffReady_ = true;
    return {};
}

Status TileRasterizer::finishFrame(const FrameParams& first, uint32_t splatCount, const RenderSettings& settings,
                                   RenderTargets& targets, FrameStats* statsOut, bool training, bool composite,
                                   const RenderTargets* under, bool scanZeroed) {
    const auto wallStart = std::chrono::steady_clock::now();
    const uint32_t tiles = first.tilesX * first.tilesY;
    uint32_t* counters = counters_.as<uint32_t>();
    const bool timed = settings.timeStages;
    FrameStats stats;
    stats.splats = splatCount;
    // Event k marks kernel boundary k (see kernelMs_): recorded only when timed;
    // event 0 was recorded before the projection.
    const auto mark = [&](int k) {
        if (timed) (void)hipEventRecord(events_[k], stream_);
    };
    mark(1);
    Status s;

    // How many pairs, how many visible and the drawn depth keys' range: one
    // scan, one readback.
    // The counters are zeroed with the look-back words: by the first
    // projection launch when it did so (scanZeroed), else by the scan.
    kern::scanExclusive(touched_.as<uint32_t>(), offsets_.as<uint32_t>(), splatCount, scanStatus_.as<uint64_t>(),
                        counters + 0, counters + 1, counters + 5, stream_, depthBlocks_.as<uint32_t>(), depthEntries_,
                        counters + 8, counters, 12, scanZeroed);
    mark(2);
    if (!(s = check(hipMemcpyAsync(hostCounters_, counters, 48, hipMemcpyDeviceToHost, stream_), "counts")).ok())
        return s;
    if (!(s = check(hipEventRecord(countsCopied_, stream_), "counts")).ok()) return s;
    // The duplication needs no count from the host: queued behind the readback
    // at the pair buffers' present capacity, it runs while the counts travel
    // and the host wakes. A frame with more pairs than that (or the first)
    // runs it again once they have grown.
    const auto duplicateInto = [&] {
        mark(3);
        kern::duplicate(first, splatCount, proj_.as<float4>(), rects_.as<uint2>(), touched_.as<uint32_t>(),
                        offsets_.as<uint32_t>(), keys_[0].as<uint64_t>(), values_[0].as<uint32_t>(),
                        training ? pairSlot_.as<uint32_t>() : nullptr, pairCapacity_, counters + 2, stream_,
                        counters + 8, counters + 10);
        mark(4);
    };
    const bool early = pairCapacity_ > 0 && (!training || pairSlot_.bytes() >= uint64_t{pairCapacity_} * 4);
    if (early) duplicateInto();
    // A frame given up here leaves nothing of its own running on the stream.
    const auto abandon = [&](Status failure) {
        (void)hipStreamSynchronize(stream_);
        return failure;
    };
    if (!(s = check(hipEventSynchronize(countsCopied_), "counts")).ok()) return abandon(s);
    const auto countsDone = std::chrono::steady_clock::now();
    const uint32_t pairs = splatCount > 0 ? hostCounters_[0] : 0u;
    const uint32_t visible = splatCount > 0 ? hostCounters_[1] : 0u;
    if (hostCounters_[5] != 0) {
        return abandon(Status::error(ErrorCode::InternalError, "the scan's look-back gave up"));
    }
    if (pairs > kMaxPairs) {
        return abandon(Status::error(ErrorCode::Unsupported, "more than 2^30 - 1 (tile, splat) pairs in one frame"));
    }
    stats.visible = visible;
    stats.pairs = pairs;
    // The sort's key: the tile over the depth key less the frame's least, in
    // as many bits as the drawn keys span (the same order as tile << 24 | key).
    kern::TileDepthLayout layout;
    if (pairs > 0) {
        const uint32_t lo = hostCounters_[8], hi = hostCounters_[9];
        if (lo > hi) return abandon(Status::error(ErrorCode::InternalError, "pairs without a drawn depth key"));
        layout.depthMin = lo;
        layout.depthBits = bitsFor(hi - lo + 1u);
    }
    const uint32_t keyBits = bitsFor(tiles) + layout.depthBits;
    stats.keyBits = keyBits;
    stats.sortPasses = kern::sortPasses(keyBits, settings.radixBits);
    // Duplicate into 64-bit keys, in slot order (again, if the early run lacked room).
    if (!early || pairs > pairCapacity_) {
        // The early run must be done before its buffers may be replaced.
        if (early && !(s = check(hipStreamSynchronize(stream_), "duplicate")).ok()) return s;
        if (!(s = reservePairs(std::max(pairs, 1u), keyBits, training, settings.radixBits)).ok()) return s;
        kern::fill32(counters + 2, 0u, 1, stream_);   // the early run's dropped pairs
        duplicateInto();
    } else if (!(s = reservePairs(std::max(pairs, 1u), keyBits, training, settings.radixBits)).ok()) {
        return abandon(s);
    }

    // One stable sort on (tile, depth >> 7), then the ranges.
    const kern::SortWorkspace ws =
        kern::sortWorkspace(sortWork_.as<uint32_t>(), pairCapacity_, keyBits, counters + 4, settings.radixBits);
    const bool inAlt = kern::sortPairs(keys_[0].as<uint64_t>(), values_[0].as<uint32_t>(), keys_[1].as<uint64_t>(),
                                       values_[1].as<uint32_t>(), pairs, kern::KeyMode::TileDepth, 0, keyBits, ws,
                                       stream_, timed ? events_[5] : nullptr, nullptr, layout,
                                       reinterpret_cast<uint32_t*>(ranges_.as<uint2>()), uint64_t{tiles} * 2,
                                       settings.radixBits);
    mark(6);
    const int sorted = inAlt ? 1 : 0;
    lastNarrow_ = kern::sortNarrows(kern::KeyMode::TileDepth, keyBits, settings.radixBits);
    lastTileShift_ = layout.depthBits - kern::narrowShift(keyBits);
    if (lastNarrow_) {
        kern::tileRangesNarrow(keys_[sorted].as<uint32_t>(), lastTileShift_, pairs, tiles, ranges_.as<uint2>(),
                               stream_, nullptr, true);
    } else {
        kern::tileRanges(keys_[sorted].as<uint64_t>(), pairs, tiles, ranges_.as<uint2>(), stream_, nullptr, true);
    }
    // A training frame's sort carried pair indices: its blend reads each
    // pair's slot through them (a separate gather of the sorted slots cost more,
    // its random reads all on the critical path).
    const uint32_t* pairSlots = values_[sorted].as<uint32_t>();
    const uint32_t* slotOfPair = training ? pairSlot_.as<uint32_t>() : nullptr;

    const uint32_t* order = nullptr;
    // Heavy tiles split across waves: frames that keep no composite or
    // training state, launched heaviest first.
    const bool split = !training && !composite && settings.heaviestTilesFirst && settings.blendSegment > 0;
    // Each pixel's last contributor and final transmittance: what a backward
    // needs, kept otherwise only on request.
    const bool keepWalk = !composite && (training || settings.keepWalkState);
    uint32_t* lastCountOut = keepWalk ? lastCount_.as<uint32_t>() : nullptr;
    float* finalTOut = keepWalk ? finalT_.as<float>() : nullptr;
    if (settings.heaviestTilesFirst && !split) {
        kern::tileOrder(ranges_.as<uint2>(), tiles, tileOrder_.as<uint32_t>(), stream_);
        order = tileOrder_.as<uint32_t>();
    }
    mark(7);   // the blend stage: the blend launch alone (a split blend's plan included)
    if (split) {
        const uint32_t seg = settings.blendSegment;
        kern::BlendPlan plan{};
        plan.taskCapacity = tiles + pairs / seg + 1u;
        plan.combineCapacity = std::min(tiles, pairs / (seg + 1u) + 1u);
        const uint64_t slotCapacity = uint64_t{pairs / seg} + plan.combineCapacity + 1u;
        const uint64_t tilePixels = uint64_t{settings.tileSize} * settings.tileSize;
        const uint64_t partialBytes = slotCapacity * tilePixels * 16u;
        if (!(s = planTasks_.reserve(uint64_t{plan.taskCapacity} * 16, "blend.tasks")).ok()) return s;
        if (!(s = planCombine_.reserve(uint64_t{plan.combineCapacity} * 16, "blend.combine")).ok()) return s;
        if (!(s = partials_.reserve(partialBytes + slotCapacity * tilePixels * 8u, "blend.partials")).ok()) return s;
        plan.order = tileOrder_.as<uint32_t>();
        plan.tasks = planTasks_.as<uint4>();
        plan.combine = planCombine_.as<uint4>();
        plan.counts = counters + 6;
        plan.partialA = partials_.as<float4>();
        plan.partialB = reinterpret_cast<float2*>(partials_.as<uint8_t>() + partialBytes);
        kern::blendSplit(first, settings.blendBatch, seg, splatCount, ranges_.as<uint2>(), pairSlots,
                         proj_.as<float4>(), plan, targets.colour, targets.depth, lastCountOut, finalTOut, stream_);
    } else if (composite) {
        kern::blendComposite(first, settings.blendBatch, splatCount, order, ranges_.as<uint2>(), pairSlots,
                             proj_.as<float4>(), under != nullptr ? under->colour : nullptr,
                             under != nullptr ? under->depth : nullptr, targets.colour, targets.depth, stream_);
    } else {
        kern::blend(first, settings.blendBatch, splatCount, order, ranges_.as<uint2>(), pairSlots,
                    proj_.as<float4>(), targets.colour, targets.depth, lastCountOut, finalTOut,
                    training ? accum_.as<float4>() : nullptr, stream_, nullptr, slotOfPair);
    }
    lastBlendPlain_ = !split && !composite && !training;   // a training frame's accum_ is kept
    lastBlendFrame_ = first;
    lastBlendList_ = settings.blendBatch;
    lastBlendSlots_ = splatCount;
    lastBlendOrder_ = order;
    lastBlendPairSlots_ = pairSlots;
    mark(8);
    if (!(s = check(hipMemcpyAsync(hostCounters_, counters, 48, hipMemcpyDeviceToHost, stream_), "frame")).ok())
        return s;
    if (!(s = check(hipStreamSynchronize(stream_), "frame")).ok()) return s;
    if (!(s = check(hipGetLastError(), "frame")).ok()) return s;
    if (hostCounters_[2] != 0) return Status::error(ErrorCode::InternalError, "the duplication overran its pairs");
    if (hostCounters_[10] != 0) {
        return Status::error(ErrorCode::InternalError, "a drawn splat's depth key fell outside the frame's range");
    }
    if (hostCounters_[4] != 0) return Status::error(ErrorCode::InternalError, "the sort's look-back gave up");
    const auto wallEnd = std::chrono::steady_clock::now();

    lastSplats_ = splatCount;
    lastVisible_ = visible;
    lastPairs_ = pairs;
    lastTiles_ = tiles;
    lastOrdered_ = settings.heaviestTilesFirst;
    lastWalkDropped_ = !keepWalk;
    lastPixels_ = uint64_t{settings.width} * settings.height;
    lastSorted_ = sorted;
    lastSlotsSorted_ = training;
    stats.totalMs = std::chrono::duration<double, std::milli>(wallEnd - wallStart).count();
    if (timed) {
        // project, scan, duplicate, sort histogram, sort passes, ranges, blend
        static constexpr int kFrom[7] = {0, 1, 3, 4, 5, 6, 7};
        for (int k = 0; k < 7; ++k) {
            float ms = 0.0F;
            (void)hipEventElapsedTime(&ms, events_[kFrom[k]], events_[kFrom[k] + 1]);
            kernelMs_[k] = ms;
        }
        kernelMs_[7] = kernelMs_[0] + kernelMs_[1] + kernelMs_[2] + kernelMs_[3] + kernelMs_[4] + kernelMs_[5] +
                       kernelMs_[6];
        stats.projectMs = kernelMs_[0];
        stats.countsMs = kernelMs_[1] + std::chrono::duration<double, std::milli>(countsDone - wallStart).count() -
                         kernelMs_[0] - kernelMs_[1];
        stats.emitMs = kernelMs_[2];
        stats.tileSortMs = kernelMs_[3] + kernelMs_[4] + kernelMs_[5];
        stats.blendMs = kernelMs_[6];
    }
    if (statsOut != nullptr) *statsOut = stats;
    return {};
}

void TileRasterizer::enqueueViews(std::span<const SplatInstance> instances, const kern::DiffCloud* diff,
                                  uint32_t viewCount, uint32_t splatsPerView, uint32_t tiles,
                                  const RenderSettings& settings, hipStream_t stream) {
    const uint32_t entries = viewCount * static_cast<uint32_t>(std::max<size_t>(instances.size(), 1));
    uint32_t* counters = counters_.as<uint32_t>();
    auto* params = viewParams_.as<FrameParams>();
    auto* views = viewInfo_.as<ViewInfo>();
    const bool training = diff != nullptr;
    // Stage events as render()'s (kernelMs_), only when not capturing.
    const bool timed = settings.timeStages && stream == stream_;
    const auto mark = [&](int k) {
        if (timed) (void)hipEventRecord(events_[k], stream);
    };
    (void)hipMemcpyAsync(params, hostParams_, sizeof(FrameParams) * entries, hipMemcpyHostToDevice, stream);
    (void)hipMemcpyAsync(views, hostViews_, sizeof(ViewInfo) * viewCount, hipMemcpyHostToDevice, stream);
    mark(0);
    // Project: one launch per instance covers every view.
    if (training) {
        kern::projectDiffViews(params, viewCount, *diff, proj_.as<float4>(), rects_.as<uint2>(),
                               touched_.as<uint32_t>(), stream);
    } else {
        for (size_t k = 0; k < instances.size(); ++k) {
            const SplatCloud* cloud = instances[k].splats;
            if (cloud == nullptr || cloud->count == 0) continue;
            kern::projectViews(params + k * viewCount, viewCount, cloud->count, cloud->positions, cloud->shape,
                               cloud->sh, cloud->filter3d, proj_.as<float4>(), rects_.as<uint2>(),
                               touched_.as<uint32_t>(), stream);
        }
    }
    const uint32_t slots = viewCount * splatsPerView;
    mark(1);
    kern::fill32(counters, 0u, 11, stream);
    kern::scanExclusive(touched_.as<uint32_t>(), offsets_.as<uint32_t>(), slots, scanStatus_.as<uint64_t>(),
                        counters + 0, counters + 1, counters + 5, stream);
    mark(2);
    mark(3);
    // Every drawn slot's depth key is checked against the sort key's field.
    kern::duplicateViews(views, splatsPerView, tileShiftOf(settings.tileSize), slots, proj_.as<float4>(), rects_.as<uint2>(),
                         touched_.as<uint32_t>(), offsets_.as<uint32_t>(), keys_[0].as<uint64_t>(),
                         values_[0].as<uint32_t>(), training ? pairSlot_.as<uint32_t>() : nullptr, pairCapacity_,
                         counters + 2, counters + 10, stream);
    mark(4);
    // The pair count (counters[0], from the scan) stays on the device: the
    // sort and the ranges read it, launched for the capacity.
    const uint32_t keyBits = bitsFor(tiles) + viewsLayout_.depthBits;
    const kern::SortWorkspace ws =
        kern::sortWorkspace(sortWork_.as<uint32_t>(), pairCapacity_, keyBits, counters + 4, settings.radixBits);
    const bool inAlt = kern::sortPairs(keys_[0].as<uint64_t>(), values_[0].as<uint32_t>(), keys_[1].as<uint64_t>(),
                                       values_[1].as<uint32_t>(), pairCapacity_, kern::KeyMode::TileDepth, 0,
                                       keyBits, ws, stream, timed ? events_[5] : nullptr, counters + 0, viewsLayout_,
                                       nullptr, 0, settings.radixBits);
    mark(6);
    const int sorted = inAlt ? 1 : 0;
    lastNarrow_ = kern::sortNarrows(kern::KeyMode::TileDepth, keyBits, settings.radixBits);
    lastTileShift_ = viewsLayout_.depthBits - kern::narrowShift(keyBits);
    if (lastNarrow_) {
        kern::tileRangesNarrow(keys_[sorted].as<uint32_t>(), lastTileShift_, pairCapacity_, tiles, ranges_.as<uint2>(),
                               stream, counters + 0);
    } else {
        kern::tileRanges(keys_[sorted].as<uint64_t>(), pairCapacity_, tiles, ranges_.as<uint2>(), stream,
                         counters + 0);
    }
    // A training batch's sort carried pair indices: the blend reads each pair's
    // slot through them.
    const uint32_t* pairSlots = values_[sorted].as<uint32_t>();
    const uint32_t* slotOfPair = training ? pairSlot_.as<uint32_t>() : nullptr;
    mark(7);
    // A persistent blend keeps about 24 one-wave workgroups on each CU.
    const bool keepWalk = training || settings.keepWalkState;
    kern::blendViews(settings.depthMode, settings.blendBatch, settings.persistentBlend, computeUnits_ * 24u, params,
                     views, viewCount, tiles, tileShiftOf(settings.tileSize), slots, ranges_.as<uint2>(), pairSlots,
                     proj_.as<float4>(),
                     keepWalk ? lastCount_.as<uint32_t>() : nullptr, keepWalk ? finalT_.as<float>() : nullptr,
                     training ? accum_.as<float4>() : nullptr, counters + 6, stream, slotOfPair);
    mark(8);
    (void)hipMemcpyAsync(hostCounters_, counters, 48, hipMemcpyDeviceToHost, stream);
    lastSorted_ = sorted;
    lastSlotsSorted_ = training;
}

Status TileRasterizer::renderViews(std::span<const View> views, std::span<const SplatInstance> instances,
                                   const RenderSettings& settings, FrameStats* statsOut) {
    for (const SplatInstance& instance : instances) {
        if (instance.splats == nullptr) continue;
        const SplatCloud& c = *instance.splats;
        if (c.count > 0 && (c.positions == nullptr || c.shape == nullptr)) {
            return Status::error(ErrorCode::InvalidArgument, "a cloud without positions or shape");
        }
        if (c.count > 0 && c.restPerColour > 0 && c.sh == nullptr) {
            return Status::error(ErrorCode::InvalidArgument, "a cloud with harmonics but no sh buffer");
        }
        if (c.restPerColour != 0 && c.restPerColour != 3 && c.restPerColour != 8 && c.restPerColour != 15) {
            return Status::error(ErrorCode::InvalidArgument, "restPerColour must be 0, 3, 8 or 15");
        }
        if (c.restPerColour > 0 && (c.shWords < (c.restPerColour * 3 + 1) / 2 || c.shWords > 23)) {
            return Status::error(ErrorCode::InvalidArgument, "shWords does not fit restPerColour");
        }
        if (settings.filter3d && c.count > 0 && c.filter3d == nullptr) {
            return Status::error(ErrorCode::InvalidArgument, "the 3D filter is on and a cloud has no filter3d");
        }
    }
    return runViews(views, instances, nullptr, settings, statsOut);
}

Status TileRasterizer::renderDiffViews(std::span<const View> views, const kern::DiffCloud& cloud,
                                       const Mat4& objectToWorld, const RenderSettings& settings,
                                       FrameStats* statsOut) {
    if (cloud.count > 0 && (cloud.means == nullptr || cloud.logScales == nullptr || cloud.quats == nullptr ||
                            cloud.logits == nullptr || cloud.sh == nullptr)) {
        return Status::error(ErrorCode::InvalidArgument, "a differentiable cloud with a missing parameter");
    }
    if (cloud.restPerColour != 0 && cloud.restPerColour != 3 && cloud.restPerColour != 8 &&
        cloud.restPerColour != 15) {
        return Status::error(ErrorCode::InvalidArgument, "restPerColour must be 0, 3, 8 or 15");
    }
    if (settings.filter3d && cloud.count > 0 && cloud.filter3d == nullptr) {
        return Status::error(ErrorCode::InvalidArgument, "the 3D filter is on and the cloud has no filter3d");
    }
    // The cloud's place in the frame constants: its count, harmonics and transform.
    SplatCloud shapeOnly;
    shapeOnly.count = cloud.count;
    shapeOnly.restPerColour = cloud.restPerColour;
    SplatInstance instance;
    instance.splats = &shapeOnly;
    instance.objectToWorld = objectToWorld;
    Status s = runViews(views, std::span<const SplatInstance>(&instance, 1), &cloud, settings, statsOut);
    if (!s.ok()) return s;
    diffCloud_ = cloud;
    diffViews_ = static_cast<uint32_t>(views.size());
    diffViewsReady_ = true;
    return {};
}

Status TileRasterizer::runViews(std::span<const View> views, std::span<const SplatInstance> instances,
                                const kern::DiffCloud* diff, const RenderSettings& settings, FrameStats* statsOut) {
    if (!created_) return Status::error(ErrorCode::InvalidArgument, "the rasteriser was not created");
    if (views.empty()) return Status::error(ErrorCode::InvalidArgument, "a batch needs a view");
    const bool training = diff != nullptr;
    if (settings.tileSize != 8 && settings.tileSize != 16 && settings.tileSize != 32) {
        return Status::error(ErrorCode::InvalidArgument, "a tile size other than 8, 16 or 32 pixels");
    }
    if (settings.radixBits != 0 && (settings.radixBits < 8 || settings.radixBits > 10)) {
        return Status::error(ErrorCode::InvalidArgument, "a sort digit other than 8, 9 or 10 bits (or 0)");
    }
    if (training && settings.tileSize != kTile) {
        return Status::error(ErrorCode::Unsupported, "a differentiable frame with tiles other than 16 pixels");
    }
    const uint32_t shift = tileShiftOf(settings.tileSize);
    uint64_t pixels = 0, tiles = 0;
    for (const View& v : views) {
        const RenderTargets& t = v.targets;
        if (t.width == 0 || t.height == 0 || t.colour == nullptr) {
            return Status::error(ErrorCode::InvalidArgument, "a view without a size or a colour target");
        }
        if (tilesAlong(t.width, shift) > 0xFFFFu || tilesAlong(t.height, shift) > 0xFFFFu) {
            return Status::error(ErrorCode::Unsupported, "a frame over 65535 tiles wide or tall");
        }
        pixels += uint64_t{t.width} * t.height;
        tiles += uint64_t{tilesAlong(t.width, shift)} * tilesAlong(t.height, shift);
    }
    uint64_t splats = 0;
    for (const SplatInstance& instance : instances) {
        if (instance.splats != nullptr) splats += instance.splats->count;
    }
    const uint64_t slots = splats * views.size();
    if (slots > 0xFFFFFFFFull || tiles > (1ull << 24) || views.size() > 0xFFFFu) {
        return Status::error(ErrorCode::Unsupported, "a batch over 2^32 slots, 2^24 tiles or 65535 views");
    }
    const uint32_t viewCount = static_cast<uint32_t>(views.size());
    const uint32_t splatsPerView = static_cast<uint32_t>(splats);
    const uint32_t entries = viewCount * static_cast<uint32_t>(std::max<size_t>(instances.size(), 1));
    Status s = check(hipSetDevice(device_), "hipSetDevice");
    if (!s.ok()) return s;
    // The scratch this batch may regrow holds the last differentiable frame.
    diffReady_ = false;
    diffViewsReady_ = false;
    lastBlendPlain_ = false;
    if (!(s = reserveSplats(static_cast<uint32_t>(std::max<uint64_t>(slots, 1)))).ok()) return s;
    if (tiles > tileCapacity_) {
        if (!(s = ranges_.reserve(tiles * 8, "tile.ranges")).ok()) return s;
        if (!(s = tileOrder_.reserve(tiles * 4, "tile.order")).ok()) return s;
        tileCapacity_ = static_cast<uint32_t>(tiles);
    }
    if (pixels > pixelCapacity_) {
        if (!(s = lastCount_.reserve(pixels * 4, "pixel.lastCount")).ok()) return s;
        if (!(s = finalT_.reserve(pixels * 4, "pixel.finalT")).ok()) return s;
        pixelCapacity_ = pixels;
    }
    if (training && !(s = accum_.reserve(pixels * 16, "pixel.accum")).ok()) return s;
    if (entries > batchCapacity_) {
        if (hostParams_ != nullptr) (void)hipHostFree(hostParams_);
        if (hostViews_ != nullptr) (void)hipHostFree(hostViews_);
        hostParams_ = nullptr;
        hostViews_ = nullptr;
        batchCapacity_ = 0;
        if (!(s = check(hipHostMalloc(reinterpret_cast<void**>(&hostParams_), sizeof(FrameParams) * entries,
                                      hipHostMallocDefault), "hipHostMalloc")).ok())
            return s;
        if (!(s = check(hipHostMalloc(reinterpret_cast<void**>(&hostViews_), sizeof(ViewInfo) * entries,
                                      hipHostMallocDefault), "hipHostMalloc")).ok())
            return s;
        if (!(s = viewParams_.reserve(sizeof(FrameParams) * entries, "batch.params")).ok()) return s;
        if (!(s = viewInfo_.reserve(sizeof(ViewInfo) * entries, "batch.views")).ok()) return s;
        batchCapacity_ = entries;
    }
    // The frame constants: entry k * views + v is instance k seen by view v,
    // its slots from v * splatsPerView + the instance's base.
    uint32_t tileBase = 0;
    uint64_t pixelBase = 0;
    // A view draws only splats with near < z <= far (as floats, as the
    // projection tests them), so the batch's depth keys lie between the keys of
    // its least near and greatest far plane: the sort's depth field, fixed
    // before any count comes back. Planes that bound nothing keep all 24 bits.
    float nearest = std::numeric_limits<float>::infinity(), farthest = 0.0f;
    bool bounded = viewCount > 0;
    for (uint32_t v = 0; v < viewCount; ++v) {
        const float nearZ = static_cast<float>(views[v].projection.nearZ);
        const float farZ = static_cast<float>(views[v].projection.farZ);
        bounded = bounded && nearZ > 0.0f && farZ >= nearZ && farZ < std::numeric_limits<float>::infinity();
        nearest = std::min(nearest, nearZ);
        farthest = std::max(farthest, farZ);
    }
    viewsLayout_ = kern::TileDepthLayout{};
    if (bounded) {
        const auto depthKey = [](float z) { return (std::bit_cast<uint32_t>(z) & 0x7FFFFFFFu) >> 7; };
        viewsLayout_.depthMin = depthKey(nearest);
        viewsLayout_.depthBits = bitsFor(depthKey(farthest) - depthKey(nearest) + 1u);
    }
    const uint32_t depthHi = viewsLayout_.depthBits >= 24
                                 ? 0x00FFFFFFu
                                 : viewsLayout_.depthMin + ((1u << viewsLayout_.depthBits) - 1u);
    for (uint32_t v = 0; v < viewCount; ++v) {
        const View& view = views[v];
        RenderSettings vs = settings;
        vs.width = view.targets.width;
        vs.height = view.targets.height;
        uint32_t base = v * splatsPerView;
        if (instances.empty()) {
            hostParams_[v] = frameParamsFor(view.projection, SplatInstance{}, vs, base);
        }
        for (size_t k = 0; k < instances.size(); ++k) {
            hostParams_[k * viewCount + v] = frameParamsFor(view.projection, instances[k], vs, base);
            if (instances[k].splats != nullptr) base += instances[k].splats->count;
        }
        ViewInfo& info = hostViews_[v];
        info.width = view.targets.width;
        info.height = view.targets.height;
        info.tilesX = tilesAlong(info.width, shift);
        info.tilesY = tilesAlong(info.height, shift);
        info.tileBase = tileBase;
        info.params = v;
        info.pixelBase = pixelBase;
        info.colour = reinterpret_cast<float*>(view.targets.colour);
        info.depth = view.targets.depth;
        info.depthLo = viewsLayout_.depthMin;
        info.depthHi = depthHi;
        tileBase += info.tilesX * info.tilesY;
        pixelBase += uint64_t{info.width} * info.height;
    }
    const uint32_t keyBits = 24 + bitsFor(static_cast<uint32_t>(tiles));
    const auto wallStart = std::chrono::steady_clock::now();
    // The pair capacity is the last batch's; a batch with more pairs overflows
    // the duplication, which the host sees after the wait: then the capacity
    // grows to the count and the batch runs again.
    for (int attempt = 0; attempt < 2; ++attempt) {
        if (!(s = reservePairs(std::max(pairCapacity_, 1u), keyBits, training, settings.radixBits)).ok()) return s;
        if (settings.useGraph) {
            // A captured graph holds its launches' sizes and every buffer
            // address: any change of either needs a new capture.
            std::string key = std::to_string(viewCount) + "/" + std::to_string(splatsPerView) + "/" +
                              std::to_string(tiles) + "/" + std::to_string(pairCapacity_) + "/" +
                              std::to_string(settings.blendBatch) + "/" + std::to_string(settings.persistentBlend) +
                              "/" + std::to_string(settings.keepWalkState) + "/" + std::to_string(training) + "/" +
                              std::to_string(settings.depthMode) + "/" + std::to_string(settings.tileSize) + "/" +
                              std::to_string(settings.radixBits) + "/" +
                              std::to_string(viewsLayout_.depthMin) +
                              "/" + std::to_string(viewsLayout_.depthBits);
            const auto addr = [&key](const void* p) { key += "/" + std::to_string(reinterpret_cast<uintptr_t>(p)); };
            for (const SplatInstance& instance : instances) {
                const SplatCloud* c = instance.splats;
                addr(c);
                if (c == nullptr) continue;
                addr(c->positions);
                addr(c->shape);
                addr(c->sh);
                addr(c->filter3d);
                key += ":" + std::to_string(c->count);
            }
            if (training) {
                for (const void* p : {static_cast<const void*>(diff->means), static_cast<const void*>(diff->logScales),
                                      static_cast<const void*>(diff->quats), static_cast<const void*>(diff->logits),
                                      static_cast<const void*>(diff->sh), static_cast<const void*>(diff->filter3d)})
                    addr(p);
                key += ":" + std::to_string(diff->count) + ":" + std::to_string(diff->restPerColour);
            }
            for (const DeviceBuffer* b : {&proj_, &rects_, &touched_, &offsets_, &scanStatus_, &keys_[0], &keys_[1],
                                          &values_[0], &values_[1], &sortWork_, &ranges_, &lastCount_, &finalT_,
                                          &counters_, &viewParams_, &viewInfo_, &pairSlot_, &sortedSlots_, &accum_})
                addr(b->get());
            addr(hostParams_);
            addr(hostViews_);
            addr(hostCounters_);
            if (graph_ == nullptr || key != graphKey_) {
                if (graph_ != nullptr) (void)hipGraphExecDestroy(graph_);
                graph_ = nullptr;
                if (captureStream_ == nullptr &&
                    !(s = check(hipStreamCreateWithFlags(&captureStream_, hipStreamNonBlocking), "capture stream"))
                         .ok())
                    return s;
                if (!(s = check(hipStreamBeginCapture(captureStream_, hipStreamCaptureModeThreadLocal),
                                "begin capture")).ok())
                    return s;
                enqueueViews(instances, diff, viewCount, splatsPerView, static_cast<uint32_t>(tiles), settings,
                             captureStream_);
                hipGraph_t graph = nullptr;
                if (!(s = check(hipStreamEndCapture(captureStream_, &graph), "end capture")).ok()) return s;
                s = check(hipGraphInstantiate(&graph_, graph, nullptr, nullptr, 0), "graph instantiate");
                (void)hipGraphDestroy(graph);
                if (!s.ok()) {
                    graph_ = nullptr;
                    return s;
                }
                graphKey_ = key;
            }
            if (!(s = check(hipGraphLaunch(graph_, stream_), "graph launch")).ok()) return s;
        } else {
            enqueueViews(instances, diff, viewCount, splatsPerView, static_cast<uint32_t>(tiles), settings, stream_);
        }
        if (!(s = check(hipStreamSynchronize(stream_), "batch")).ok()) return s;
        if (!(s = check(hipGetLastError(), "batch")).ok()) return s;
        if (hostCounters_[5] != 0) return Status::error(ErrorCode::InternalError, "the scan's look-back gave up");
        const uint32_t pairs = hostCounters_[0];
        if (pairs > kMaxPairs) {
            return Status::error(ErrorCode::Unsupported, "more than 2^30 - 1 (tile, splat) pairs in one batch");
        }
        if (hostCounters_[2] == 0) break;
        if (attempt == 1) return Status::error(ErrorCode::InternalError, "the duplication overran its pairs");
        if (!(s = reservePairs(pairs, keyBits, training, settings.radixBits)).ok()) return s;
    }
    if (hostCounters_[10] != 0) {
        return Status::error(ErrorCode::InternalError, "a drawn splat's depth key fell outside the batch's planes");
    }
    if (hostCounters_[4] != 0) return Status::error(ErrorCode::InternalError, "the sort's look-back gave up");
    const auto wallEnd = std::chrono::steady_clock::now();
    if (settings.timeStages && !settings.useGraph) {
        static constexpr int kFrom[7] = {0, 1, 3, 4, 5, 6, 7};
        float sum = 0.0F;
        for (int k = 0; k < 7; ++k) {
            float ms = 0.0F;
            (void)hipEventElapsedTime(&ms, events_[kFrom[k]], events_[kFrom[k] + 1]);
            kernelMs_[k] = ms;
            sum += ms;
        }
        kernelMs_[7] = sum;
    }
    lastSplats_ = static_cast<uint32_t>(slots);
    lastVisible_ = hostCounters_[1];
    lastPairs_ = hostCounters_[0];
    lastTiles_ = static_cast<uint32_t>(tiles);
    lastOrdered_ = false;
    lastWalkDropped_ = !(training || settings.keepWalkState);
    lastPixels_ = pixels;
    if (statsOut != nullptr) {
        FrameStats stats;
        stats.splats = static_cast<uint32_t>(slots);
        stats.visible = lastVisible_;
        stats.pairs = lastPairs_;
        stats.keyBits = bitsFor(static_cast<uint32_t>(tiles)) + viewsLayout_.depthBits;
        stats.sortPasses = kern::sortPasses(stats.keyBits, settings.radixBits);
        stats.totalMs = std::chrono::duration<double, std::milli>(wallEnd - wallStart).count();
        *statsOut = stats;
    }
    return {};
}

Status TileRasterizer::referenceRender(const Projection& projection, std::span<const SplatInstance> instances,
                                       const RenderSettings& settings, double* colour, double* depth, bool keyed,
                                       const uint32_t* pixels, uint32_t pixelCount) {
    if (!created_) return Status::error(ErrorCode::InvalidArgument, "the rasteriser was not created");
    if (settings.width == 0 || settings.height == 0 || colour == nullptr) {
        return Status::error(ErrorCode::InvalidArgument, "a frame needs a size and a colour target");
    }
    uint64_t total = 0;
    for (const SplatInstance& instance : instances) {
        if (instance.splats == nullptr) continue;
        const SplatCloud& c = *instance.splats;
        if (c.count > 0 && (c.positions == nullptr || c.shape == nullptr)) {
            return Status::error(ErrorCode::InvalidArgument, "a cloud without positions or shape");
        }
        if (c.count > 0 && c.restPerColour > 0 && c.sh == nullptr) {
            return Status::error(ErrorCode::InvalidArgument, "a cloud with harmonics but no sh buffer");
        }
        if (settings.filter3d && c.count > 0 && c.filter3d == nullptr) {
            return Status::error(ErrorCode::InvalidArgument, "the 3D filter is on and a cloud has no filter3d");
        }
        total += c.count;
    }
    if (total > kMaxPairs) return Status::error(ErrorCode::Unsupported, "more than 2^30 - 1 splats");
    const uint32_t slots = static_cast<uint32_t>(std::max<uint64_t>(total, 1));
    Status s = check(hipSetDevice(device_), "hipSetDevice");
    if (!s.ok()) return s;
    if (!(s = refRecords_.reserve(uint64_t{slots} * kern::referenceRecordBytes(), "reference.records")).ok()) return s;
    for (int k = 0; k < 2; ++k) {
        if (!(s = refKeys_[k].reserve(uint64_t{slots} * 8, "reference.keys")).ok()) return s;
        if (!(s = refValues_[k].reserve(uint64_t{slots} * 4, "reference.values")).ok()) return s;
    }
    const uint64_t words = kern::sortWorkspaceWords(slots, 64);
    if (!(s = refWork_.reserve(words * 4, "reference.sortWorkspace")).ok()) return s;
    uint32_t* counters = counters_.as<uint32_t>();
    kern::fill32(counters, 0u, 8, stream_);
    uint32_t base = 0;
    FrameParams first{};
    bool haveFirst = false;
    for (const SplatInstance& instance : instances) {
        const SplatCloud* cloud = instance.splats;
        const FrameParams f = frameParamsFor(projection, instance, settings, base);
        if (!haveFirst) {
            first = f;
            haveFirst = true;
        }
        if (cloud == nullptr || cloud->count == 0) continue;
        kern::referenceProject(f, cloud->positions, cloud->shape, cloud->sh, cloud->filter3d, refRecords_.get(),
                               refKeys_[0].as<uint64_t>(), refValues_[0].as<uint32_t>(), counters + 7, keyed, stream_);
        base += cloud->count;
    }
    if (!haveFirst) first = frameParamsFor(projection, SplatInstance{}, settings, 0);
    const kern::SortWorkspace ws =
        kern::sortWorkspace(refWork_.as<uint32_t>(), static_cast<uint32_t>(total), 64, counters + 4);
    const bool alt = kern::sortPairs(refKeys_[0].as<uint64_t>(), refValues_[0].as<uint32_t>(),
                                     refKeys_[1].as<uint64_t>(), refValues_[1].as<uint32_t>(),
                                     static_cast<uint32_t>(total), kern::KeyMode::Plain, 0, 64, ws, stream_);
    kern::referenceBlend(first, refRecords_.get(), refValues_[alt ? 1 : 0].as<uint32_t>(), counters + 7, colour,
                         depth, stream_, pixels, pixels != nullptr ? pixelCount : 0u);
    if (!(s = check(hipMemcpyAsync(hostCounters_, counters, 32, hipMemcpyDeviceToHost, stream_), "reference")).ok())
        return s;
    if (!(s = check(hipStreamSynchronize(stream_), "reference")).ok()) return s;
    if (!(s = check(hipGetLastError(), "reference")).ok()) return s;
    if (hostCounters_[4] != 0) return Status::error(ErrorCode::InternalError, "the sort's look-back gave up");
    return {};
}

Status TileRasterizer::autotuneViews(std::span<const View> views, std::span<const SplatInstance> instances,
                                     RenderSettings& settings, float* ms) {
    RenderSettings best = settings;
    double bestMs = -1.0;
    for (const uint32_t tile : {8u, 16u, 32u})
    for (const uint32_t list : {64u, 128u, 256u}) {
        for (const bool persistent : {false, true}) {
            for (const bool graph : {false, true}) {
                RenderSettings trial = settings;
                trial.tileSize = tile;
                trial.blendBatch = list;
                trial.persistentBlend = persistent;
                trial.useGraph = graph;
                Status s = renderViews(views, instances, trial, nullptr);   // sizes buffers, captures
                if (!s.ok()) return s;
                double fastest = -1.0;
                for (int rep = 0; rep < 3; ++rep) {
                    const auto t0 = std::chrono::steady_clock::now();
                    if (!(s = renderViews(views, instances, trial, nullptr)).ok()) return s;
                    const double t = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
                                         .count();
                    if (fastest < 0.0 || t < fastest) fastest = t;
                }
                if (bestMs < 0.0 || fastest < bestMs) {
                    bestMs = fastest;
                    best = trial;
                }
            }
        }
    }
    // The sort's digit width under the fastest launch choices.
    const RenderSettings launch = best;
    for (const uint32_t radix : {8u, 9u, 10u}) {
        RenderSettings trial = launch;
        trial.radixBits = radix;
        Status s = renderViews(views, instances, trial, nullptr);
        if (!s.ok()) return s;
        double fastest = -1.0;
        for (int rep = 0; rep < 3; ++rep) {
            const auto t0 = std::chrono::steady_clock::now();
            if (!(s = renderViews(views, instances, trial, nullptr)).ok()) return s;
            const double t = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            if (fastest < 0.0 || t < fastest) fastest = t;
        }
        if (fastest < bestMs) {
            bestMs = fastest;
            best = trial;
        }
    }
    settings = best;
    if (ms != nullptr) *ms = static_cast<float>(bestMs);
    return {};
}

Status TileRasterizer::backward(const float4* gradColour, const kern::DiffGrads& grads, float* millis) {
    if (!created_) return Status::error(ErrorCode::InvalidArgument, "the rasteriser was not created");
    if (!diffReady_) {
        return Status::error(ErrorCode::InvalidArgument, "no differentiable render to take the backward of");
    }
    if (gradColour == nullptr) return Status::error(ErrorCode::InvalidArgument, "no colour gradient");
    if (diffCloud_.count > 0 && (grads.means == nullptr || grads.logScales == nullptr || grads.quats == nullptr ||
                                 grads.logits == nullptr || grads.sh == nullptr)) {
        return Status::error(ErrorCode::InvalidArgument, "a missing gradient buffer");
    }
    Status s = check(hipSetDevice(device_), "hipSetDevice");
    if (!s.ok()) return s;
    if (!(s = pairGrad_.reserve(uint64_t{std::max(pairCapacity_, 1u)} * kern::kPairGradBytes, "pair.grad")).ok()) return s;
    const kern::PairGrads pairGrads = kern::pairGradsIn(pairGrad_.get(), std::max(pairCapacity_, 1u));
    if (millis != nullptr) (void)hipEventRecord(events_[0], stream_);
    // One view's tiles in raster order: neighbouring tiles share splats in L2,
    // and on a 1M-splat 720p frame that beat heaviest-first by about 10%.
    kern::blendBackward(diffFrame_, lastSplats_, ranges_.as<uint2>(), pairSlot_.as<uint32_t>(),
                        values_[lastSorted_].as<uint32_t>(), proj_.as<float4>(), accum_.as<float4>(),
                        lastCount_.as<uint32_t>(), finalT_.as<float>(), gradColour, pairGrads, stream_,
                        nullptr);
    if (millis != nullptr) (void)hipEventRecord(events_[1], stream_);
    kern::projectBackward(diffFrame_, diffCloud_, touched_.as<uint32_t>(), offsets_.as<uint32_t>(),
                          pairGrads, grads, stream_);
    if (millis != nullptr) (void)hipEventRecord(events_[2], stream_);
    if (!(s = check(hipStreamSynchronize(stream_), "backward")).ok()) return s;
    if (!(s = check(hipGetLastError(), "backward")).ok()) return s;
    if (millis != nullptr) {
        (void)hipEventElapsedTime(&millis[0], events_[0], events_[1]);
        (void)hipEventElapsedTime(&millis[1], events_[1], events_[2]);
    }
    return {};
}

Status TileRasterizer::backwardViews(std::span<const float4* const> gradColours, const kern::DiffGrads& grads,
                                     float* millis) {
    if (!created_) return Status::error(ErrorCode::InvalidArgument, "the rasteriser was not created");
    if (!diffViewsReady_) {
        return Status::error(ErrorCode::InvalidArgument, "no differentiable batch to take the backward of");
    }
    if (gradColours.size() != diffViews_) {
        return Status::error(ErrorCode::InvalidArgument, "one colour gradient a view of the batch is needed");
    }
    for (const float4* g : gradColours) {
        if (g == nullptr) return Status::error(ErrorCode::InvalidArgument, "no colour gradient");
    }
    if (diffCloud_.count > 0 && (grads.means == nullptr || grads.logScales == nullptr || grads.quats == nullptr ||
                                 grads.logits == nullptr || grads.sh == nullptr)) {
        return Status::error(ErrorCode::InvalidArgument, "a missing gradient buffer");
    }
    Status s = check(hipSetDevice(device_), "hipSetDevice");
    if (!s.ok()) return s;
    if (!(s = pairGrad_.reserve(uint64_t{std::max(pairCapacity_, 1u)} * kern::kPairGradBytes, "pair.grad")).ok()) return s;
    const kern::PairGrads pairGrads = kern::pairGradsIn(pairGrad_.get(), std::max(pairCapacity_, 1u));
    if (diffViews_ > gradPtrCapacity_) {
        if (hostGradPtrs_ != nullptr) (void)hipHostFree(hostGradPtrs_);
        hostGradPtrs_ = nullptr;
        gradPtrCapacity_ = 0;
        if (!(s = check(hipHostMalloc(reinterpret_cast<void**>(&hostGradPtrs_), sizeof(float4*) * diffViews_,
                                      hipHostMallocDefault), "hipHostMalloc")).ok())
            return s;
        if (!(s = viewGrads_.reserve(sizeof(float4*) * diffViews_, "batch.gradColours")).ok()) return s;
        gradPtrCapacity_ = diffViews_;
    }
    // The pinned table is rewritten only once the last backward's copy of it is done.
    if (!(s = check(hipStreamSynchronize(stream_), "backward")).ok()) return s;
    for (uint32_t v = 0; v < diffViews_; ++v) hostGradPtrs_[v] = gradColours[v];
    if (!(s = check(hipMemcpyAsync(viewGrads_.get(), hostGradPtrs_, sizeof(float4*) * diffViews_,
                                   hipMemcpyHostToDevice, stream_), "gradient table")).ok())
        return s;
    const auto* params = viewParams_.as<FrameParams>();
    if (millis != nullptr) (void)hipEventRecord(events_[0], stream_);
    // The batch's tiles heaviest first: a view of few, long tiles (a small view
    // close to the cloud) would otherwise finish alone at the end.
    kern::tileOrder(ranges_.as<uint2>(), lastTiles_, tileOrder_.as<uint32_t>(), stream_);
    kern::blendBackwardViews(params, viewInfo_.as<ViewInfo>(), diffViews_, lastTiles_, lastSplats_, ranges_.as<uint2>(),
                             pairSlot_.as<uint32_t>(), values_[lastSorted_].as<uint32_t>(), proj_.as<float4>(),
                             accum_.as<float4>(), lastCount_.as<uint32_t>(), finalT_.as<float>(),
                             viewGrads_.as<const float4*>(), pairGrads, stream_, tileOrder_.as<uint32_t>());
    if (millis != nullptr) (void)hipEventRecord(events_[1], stream_);
    kern::projectBackwardViews(params, diffViews_, diffCloud_, touched_.as<uint32_t>(), offsets_.as<uint32_t>(),
                               pairGrads, grads, stream_);
    if (millis != nullptr) (void)hipEventRecord(events_[2], stream_);
    if (!(s = check(hipStreamSynchronize(stream_), "backward")).ok()) return s;
    if (!(s = check(hipGetLastError(), "backward")).ok()) return s;
    if (millis != nullptr) {
        (void)hipEventElapsedTime(&millis[0], events_[0], events_[1]);
        (void)hipEventElapsedTime(&millis[1], events_[1], events_[2]);
    }
    return {};
}

Status TileRasterizer::prefixSum(const uint32_t* in, uint32_t* out, uint32_t* total, uint32_t count) {
    if (!created_) return Status::error(ErrorCode::InvalidArgument, "the rasteriser was not created");
    diffReady_ = false;   // the scratch it may regrow holds the last frame
    diffViewsReady_ = false;
    lastBlendPlain_ = false;
    if (in == nullptr || out == nullptr || total == nullptr) return Status::error(ErrorCode::InvalidArgument, "null buffer");
    if (in == out) return Status::error(ErrorCode::InvalidArgument, "the scan's input and output must differ");
    Status s = check(hipSetDevice(device_), "hipSetDevice");
    if (!s.ok()) return s;
    if (!(s = reserveSplats(std::max(count, 1u))).ok()) return s;
    uint32_t* counters = counters_.as<uint32_t>();
    kern::scanExclusive(in, out, count, scanStatus_.as<uint64_t>(), total, counters + 1, counters + 5,
                        stream_);
    return check(hipGetLastError(), "prefix sum");
}

uint32_t TileRasterizer::keptPerColour(const io::SplatEncoding& e, uint32_t maxDegree) {
    static constexpr uint32_t kPerDegree[4] = {0, 3, 8, 15};
    return std::min(e.restPerColour, kPerDegree[std::min(maxDegree, 3u)]);
}

Status TileRasterizer::decodeCloud(const io::RawSplats& raw, uint32_t maxDegree, float* positions, uint32_t* shape,
                                   uint32_t* sh, CloudInfo* info) {
    if (!created_) return Status::error(ErrorCode::InvalidArgument, "the rasteriser was not created");
    const io::SplatEncoding& e = raw.encoding;
    if (raw.count == 0 || e.floatsPerRecord == 0 || raw.records.size() < uint64_t{raw.count} * e.floatsPerRecord) {
        return Status::error(ErrorCode::InvalidArgument, "'" + raw.source + "': no splat records");
    }
    const uint32_t fields[] = {e.x,    e.y,    e.z,    e.opacity, e.scale0, e.scale1, e.scale2,
                               e.rotW, e.rotX, e.rotY, e.rotZ,    e.dc0,    e.dc1,    e.dc2};
    for (uint32_t f : fields) {
        if (f >= e.floatsPerRecord) return Status::error(ErrorCode::InvalidArgument, "an encoding field past the record");
    }
    const uint32_t keep = keptPerColour(e, maxDegree);
    const uint32_t perColourInFile = e.restColourOuter != 0 ? e.filePerColour : keep;
    if (keep > 0) {
        // The last rest field the decode reads: blue's last kept basis.
        const uint64_t last = e.restColourOuter != 0 ? uint64_t{e.restBase} + 2ull * perColourInFile + (keep - 1)
                                                     : uint64_t{e.restBase} + 3ull * (keep - 1) + 2;
        if (last >= e.floatsPerRecord || perColourInFile < keep) {
            return Status::error(ErrorCode::InvalidArgument, "the rest coefficients run past the record");
        }
    }
    if (positions == nullptr || shape == nullptr || (keep > 0 && sh == nullptr)) {
        return Status::error(ErrorCode::InvalidArgument, "null output buffer");
    }
    Status s = check(hipSetDevice(device_), "hipSetDevice");
    if (!s.ok()) return s;
    diffReady_ = false;   // the per-splat scratch it reuses holds the last frame
    diffViewsReady_ = false;
    lastBlendPlain_ = false;
    lastSplats_ = lastVisible_ = lastPairs_ = lastTiles_ = 0;
    const uint64_t recordBytes = uint64_t{e.floatsPerRecord} * 4;
    constexpr uint64_t kSliceBytes = 256ull << 20;
    const auto perSlice = static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>(kSliceBytes / recordBytes, raw.count)));
    if (!(s = reserveSplats(perSlice)).ok()) return s;
    if (!(s = loadRaw_.reserve(uint64_t{perSlice} * recordBytes, "load.raw")).ok()) return s;
    if (!(s = loadBox_.reserve((uint64_t{kern::boundsPartials(raw.count)} * 6 + 8) * 4, "load.box")).ok()) return s;
    kern::DecodeParams p{};
    p.stride = e.floatsPerRecord;
    p.keep = keep;
    p.shWords = keep == 0 ? 1u : (keep * 3 + 1) / 2;
    p.x = e.x; p.y = e.y; p.z = e.z; p.opacity = e.opacity;
    p.scale0 = e.scale0; p.scale1 = e.scale1; p.scale2 = e.scale2;
    p.rotW = e.rotW; p.rotX = e.rotX; p.rotY = e.rotY; p.rotZ = e.rotZ;
    p.dc0 = e.dc0; p.dc1 = e.dc1; p.dc2 = e.dc2;
    p.restBase = e.restBase;
    p.filePerColour = perColourInFile;
    p.restColourOuter = e.restColourOuter;
    p.opacityMode = static_cast<uint32_t>(e.opacityMode);
    p.scaleMode = static_cast<uint32_t>(e.scaleMode);
    p.colourMode = static_cast<uint32_t>(e.colourMode);
    p.rotationMode = static_cast<uint32_t>(e.rotationMode);
    p.restMode = static_cast<uint32_t>(e.restMode);
    p.positionScale = 1.0f / static_cast<float>(1u << std::min(e.fractionalBits, 24u));   // exact: a power of two
    p.flipYZ = e.flipYZ;
    uint32_t* counters = counters_.as<uint32_t>();
    uint32_t written = 0;
    float uploadMs = 0.0F, decodeMs = 0.0F;
    for (uint32_t first = 0; first < raw.count; first += perSlice) {
        p.count = std::min(perSlice, raw.count - first);
        (void)hipEventRecord(events_[0], stream_);
        if (!(s = check(hipMemcpyAsync(loadRaw_.get(), raw.records.data() + uint64_t{first} * e.floatsPerRecord,
                                       uint64_t{p.count} * recordBytes, hipMemcpyHostToDevice, stream_),
                        "upload raw splats"))
                 .ok())
            return s;
        (void)hipEventRecord(events_[1], stream_);
        kern::fill32(counters, 0u, 8, stream_);
        kern::decodeSplats(loadRaw_.as<float>(), p, touched_.as<uint32_t>(), offsets_.as<uint32_t>(),
                           scanStatus_.as<uint64_t>(), counters + 0, counters + 1, counters + 5, written,
                           reinterpret_cast<float4*>(positions), reinterpret_cast<uint4*>(shape), sh, stream_);
        (void)hipEventRecord(events_[2], stream_);
        (void)hipMemcpyAsync(hostCounters_, counters, 32, hipMemcpyDeviceToHost, stream_);
        if (!(s = check(hipStreamSynchronize(stream_), "decode splats")).ok()) return s;
        if (hostCounters_[5] != 0) return Status::error(ErrorCode::InternalError, "the scan's look-back gave up");
        written += hostCounters_[0];
        float ms = 0.0F;
        if (hipEventElapsedTime(&ms, events_[0], events_[1]) == hipSuccess) uploadMs += ms;
        if (hipEventElapsedTime(&ms, events_[1], events_[2]) == hipSuccess) decodeMs += ms;
    }
    CloudInfo out;
    out.uploadMs = uploadMs;
    out.decodeMs = decodeMs;
    out.declared = raw.count;
    out.count = written;
    out.restPerColour = keep;
    out.shWords = p.shWords;
    if (written == 0) return Status::error(ErrorCode::InvalidArgument, "'" + raw.source + "': no splat survived validation");
    float* box = loadBox_.as<float>();
    kern::splatBounds(reinterpret_cast<const float4*>(positions), written, box + 8, box, stream_);
    float host[6];
    if (!(s = check(hipMemcpyAsync(host, box, sizeof host, hipMemcpyDeviceToHost, stream_), "read bounds")).ok()) return s;
    if (!(s = check(hipStreamSynchronize(stream_), "bounds")).ok()) return s;
    for (int k = 0; k < 3; ++k) {
        out.boundsMin[k] = host[k];
        out.boundsMax[k] = host[3 + k];
    }
    if (info != nullptr) *info = out;
    return {};
}

Status TileRasterizer::radixSort(uint64_t* keys, uint32_t* values, uint64_t* scratchKeys, uint32_t* scratchValues,
                                 uint32_t count, uint32_t keyBits, bool* resultInScratch, uint32_t radixBits) {
    if (!created_) return Status::error(ErrorCode::InvalidArgument, "the rasteriser was not created");
    if (keyBits < 1 || keyBits > 64) return Status::error(ErrorCode::InvalidArgument, "keyBits must be 1..64");
    if (radixBits < 8 || radixBits > 10) return Status::error(ErrorCode::InvalidArgument, "radixBits must be 8..10");
    if (count > kMaxPairs) return Status::error(ErrorCode::Unsupported, "more than 2^30 - 1 pairs");
    if (count > 0 && (keys == nullptr || values == nullptr || scratchKeys == nullptr || scratchValues == nullptr)) {
        return Status::error(ErrorCode::InvalidArgument, "null buffer");
    }
    Status s = check(hipSetDevice(device_), "hipSetDevice");
    if (!s.ok()) return s;
    const uint64_t words = kern::sortWorkspaceWords(count, keyBits, radixBits);
    if (words > sortWorkWords_) {
        if (!(s = sortWork_.reserve(words * 4, "pair.sortWorkspace")).ok()) return s;
        sortWorkWords_ = words;
    }
    const kern::SortWorkspace ws =
        kern::sortWorkspace(sortWork_.as<uint32_t>(), count, keyBits, counters_.as<uint32_t>() + 4, radixBits);
    const bool alt = kern::sortPairs(keys, values, scratchKeys, scratchValues, count, kern::KeyMode::Plain, 0, keyBits,
                                     ws, stream_, nullptr, nullptr, {}, nullptr, 0, radixBits);
    if (resultInScratch != nullptr) *resultInScratch = alt;
    return check(hipGetLastError(), "radix sort");
}

Status TileRasterizer::debugCopy(int which, void* host, size_t bytes, size_t* needed) {
    if (!created_) return Status::error(ErrorCode::InvalidArgument, "the rasteriser was not created");
    const void* src = nullptr;
    size_t size = 0;
    uint32_t counts[4] = {lastSplats_, lastVisible_, lastPairs_, lastTiles_};
    switch (which) {
    case 0: size = sizeof counts; break;
    case 1: src = touched_.get(); size = size_t{lastSplats_} * 4; break;
    case 2: src = offsets_.get(); size = size_t{lastSplats_} * 4; break;
    case 3: src = proj_.get(); size = size_t{lastSplats_} * sizeof(ProjRecord); break;
    case 4: src = keys_[lastSorted_].get(); size = size_t{lastPairs_} * 8; break;   // widened below when narrow
    case 5: src = values_[lastSorted_].get(); size = size_t{lastPairs_} * 4; break;
    case 6: src = ranges_.get(); size = size_t{lastTiles_} * 8; break;
    case 7: src = lastCount_.get(); size = lastWalkDropped_ ? 0 : size_t{lastPixels_} * 4; break;
    case 8: src = finalT_.get(); size = lastWalkDropped_ ? 0 : size_t{lastPixels_} * 4; break;
    case 9: src = rects_.get(); size = size_t{lastSplats_} * 8; break;
    case 10: size = sizeof kernelMs_; break;
    case 11: src = tileOrder_.get(); size = lastOrdered_ ? size_t{lastTiles_} * 4 : 0; break;
    case 12: size = lastBlendPlain_ ? 4 : 0; break;
    default: return Status::error(ErrorCode::InvalidArgument, "unknown debug buffer " + std::to_string(which));
    }
    if (needed != nullptr) *needed = size;
    if (bytes == 0 || host == nullptr) return {};
    if (bytes < size) return Status::error(ErrorCode::InvalidArgument, "the host buffer is too small");
    if (which == 0) {
        std::memcpy(host, counts, size);
        return {};
    }
    if (which == 10) {
        std::memcpy(host, kernelMs_, size);
        return {};
    }
    if (size == 0) return {};
    Status s = check(hipSetDevice(device_), "hipSetDevice");
    if (!s.ok()) return s;
    if (which == 12) {
        // The same blend again, its images into scratch, counting the walk.
        const uint64_t pixels = uint64_t{lastBlendFrame_.width} * lastBlendFrame_.height;
        if (!(s = accum_.reserve(std::max<uint64_t>(pixels, 1) * 16, "walk.colour")).ok()) return s;
        if (!(s = walkSteps_.reserve(4, "walk.steps")).ok()) return s;
        kern::fill32(walkSteps_.as<uint32_t>(), 0u, 1, stream_);
        kern::blend(lastBlendFrame_, lastBlendList_, lastBlendSlots_, lastBlendOrder_, ranges_.as<uint2>(),
                    lastBlendPairSlots_, proj_.as<float4>(), accum_.as<float4>(), nullptr, nullptr, nullptr, nullptr,
                    stream_, walkSteps_.as<uint32_t>());
        src = walkSteps_.get();
    }
    if (which == 4 && lastNarrow_) {
        // The sort left u32 keys: rebuild tile << 32 | bits(view z) from them and
        // the pairs' slots into the other key buffer (free once a frame is done).
        const uint32_t* slots = values_[lastSorted_].as<uint32_t>();
        if (lastSlotsSorted_) {
            // A training frame's sorted values are pair indices: their slots first.
            kern::gatherSlots(slots, pairSlot_.as<uint32_t>(), lastPairs_, sortedSlots_.as<uint32_t>(), stream_);
            slots = sortedSlots_.as<uint32_t>();
        }
        kern::widenKeys(keys_[lastSorted_].as<uint32_t>(), lastTileShift_, slots, proj_.as<float4>(), lastPairs_,
                        keys_[1 - lastSorted_].as<uint64_t>(), stream_);
        src = keys_[1 - lastSorted_].get();
    }
    if (!(s = check(hipMemcpyAsync(host, src, size, hipMemcpyDeviceToHost, stream_), "debug copy")).ok()) return s;
    return check(hipStreamSynchronize(stream_), "debug copy");
}

}   // namespace lrt
