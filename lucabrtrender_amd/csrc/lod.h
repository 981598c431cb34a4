// Levels of detail for a splat cloud, built and cut on the GPU: the shape of
// the reference's lrt::lod (modules/lod/include/lrt/lod/Lod.h:36-131) for a
// cloud held whole in device memory (the streamed store, its chunks and the
// .lrtc file are not carried).
//
// Build, once per cloud: the splats in Morton order of their position in the
// cloud's bounding cube; the octree cell of level r holding a splat is its
// code's top 3r bits, so every cell's splats are one run of that order. Each
// stored level (coarsest .. finest) has a group per non-empty cell, and each
// group one merged Gaussian from additive moments of the splats under it
// (Kerbl et al. 2024): W = sum w, S1 = sum w mu, S2 = sum w (Sigma + mu mu^T),
// C = sum w colour with w = opacity x the area of the two longest axes.
//
// Cut, every frame: a group is drawn when its cell projects to at most
// `threshold` pixels and its parent's does not; a splat when the finest
// level's cell holding it does not. Every place is drawn at exactly one level.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <memory>
#include <optional>
#include <span>
#include <vector>

#include "rasterizer.h"
#include "status.h"

namespace lrt::lod {

constexpr uint32_t kLevels = 10;        // 30-bit Morton codes
constexpr uint32_t kMomentsHead = 13;   // W, S1 (3), S2 (6), base colour (3)

struct LodBuildSettings {   // Lod.h:78-87
    float maxGroupFraction = 0.5F;   // the finest stored level's cells hold >= 1 / this splats on average
    uint32_t coarsestLevel = 1;
};

// One stored level: its merged Gaussians are [first, first + count) of the
// cloud's merged arrays.
struct LodLevel {
    uint32_t level = 0, first = 0, count = 0;
};

// A cloud with its levels of detail (Lod.h:51-76, held whole on the device).
// Owns its device memory.
struct LodCloud {
    uint32_t count = 0, restPerColour = 0, shWords = 1;
    DeviceBuffer positions, shape, sh;   // the splats in Morton order
    DeviceBuffer groups;                 // u32 a splat: its finest-level group
    DeviceBuffer starts;                 // u32 a finest-level group: its first splat
    DeviceBuffer mergedPositions, mergedShape, mergedSh;   // every level's Gaussians, coarsest first
    DeviceBuffer cells;                  // u32 a merged Gaussian: its cell's code at its level
    std::vector<LodLevel> levels;        // coarsest first, consecutive levels, never empty
    float boundsLo[3] = {0, 0, 0};
    float extent = 0.0F;                 // the bounds' largest edge, widened by 1e-4
    uint32_t merged = 0;                 // merged Gaussians over all levels

    SplatCloud splats() const;           // the Morton-ordered splats
    SplatCloud gaussians(size_t level) const;
};

// Builds the levels of `cloud` on `stream` (Lod.cpp:81-283). Synchronous.
Status build(int device, hipStream_t stream, const SplatCloud& cloud, const LodBuildSettings& settings,
             LodCloud& out);

struct LodInstance {   // Lod.h:101-106
    const LodCloud* cloud = nullptr;
    Mat4 objectToWorld;
    SplatEditParams edit;
    std::optional<float> threshold;   // its own, over the call's
};

struct CutStats {   // Lod.h:108-117, without the streaming needs
    uint32_t splats = 0;      // the cloud's own splats drawn
    uint32_t merged = 0;      // merged Gaussians drawn
    uint32_t available = 0;   // splats in the cloud
};

// CutSelector (Lod.h:119-143, Lod.cpp:350-547): each instance's cut for a view
// as splat instances the rasteriser draws. The clouds they point at belong to
// the selector and are rewritten by its next call; the gathers are queued on
// `stream` (a render on the same stream sees them), after one wait for the
// drawn counts.
class CutSelector {
public:
    CutSelector();
    ~CutSelector();
    CutSelector(const CutSelector&) = delete;
    CutSelector& operator=(const CutSelector&) = delete;

    Status select(int device, hipStream_t stream, const Projection& projection,
                  std::span<const LodInstance> instances, float threshold, std::vector<SplatInstance>& out,
                  std::vector<CutStats>* stats);

private:
    struct Frame;
    std::vector<std::unique_ptr<Frame>> frames_;
    uint32_t* hostCounts_ = nullptr;   // pinned: drawn, merged drawn, scan error
};

}   // namespace lrt::lod
