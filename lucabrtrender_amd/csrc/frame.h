// The per-launch constant block of the splat kernels and the records they pass
// between stages. Plain data, shared by the C++20 host and the HIP kernels.
//
// FrameParams carries the fields of the reference's FrameParams
// (shaders/lrt/splat/frame.slang:33-81) that the splat-only path reads, plus
// two filter parameters that the reference fixes: `dilation` (its +0.3 px^2,
// splat_project.slang:122-125) and `useFilter3d` (Mip-Splatting's per-splat 3D
// filter, which the reference does not have).
#pragma once

#include <cstdint>

namespace lrt {

inline constexpr uint32_t kTile = 16;   // frame.slang:31

// A tile size's shift (8, 16 or 32 pixels; anything else is refused before a
// frame is set up) and a frame's tiles along a side of `pixels`.
inline constexpr uint32_t tileShiftOf(uint32_t tileSize) { return tileSize == 8 ? 3u : tileSize == 32 ? 5u : 4u; }
inline constexpr uint32_t tilesAlong(uint32_t pixels, uint32_t shift) { return (pixels + (1u << shift) - 1u) >> shift; }

struct SplatEditParams {   // shaders/lrt/common/edit.slang:23-36
    uint32_t active = 0, shape = 0, mode = 2, invert = 0;
    float centre[3] = {0, 0, 0};
    float size[3] = {1, 1, 1};
    float tint[3] = {1, 1, 1};
    float saturation = 1.0F, brightness = 1.0F, opacity = 1.0F, minOpacity = 0.0F, maxScale = 0.0F;
};

struct FrameParams {
    uint32_t width, height, tilesX, tilesY;
    uint32_t tileShift;   // tiles are 1 << tileShift pixels square: 3, 4 (the reference's 16) or 5
    float focalX, focalY, centreX, centreY, nearZ, farZ;
    // Where the Jacobian is linearised at most: 1.3x the frustum, in units of
    // view x / z and y / z (set with the lens; the same float operations the
    // preprocess would make a splat).
    float limLoX, limHiX, limLoY, limHiY;
    uint32_t orthographic, antialias;
    float m[12];          // object -> view, rows of a 3x4
    float eye[3];         // the eye in the cloud's own space
    uint32_t count;       // splats of this instance
    uint32_t base;        // its first slot
    uint32_t restPerColour, shWords, shLimit, tileBits;
    uint32_t depthMode;
    float depthThreshold;
    float bg[4];
    uint32_t linearise;
    float dilation;
    uint32_t useFilter3d;
    SplatEditParams edit;
};

// How a point cloud's discs are sized and coloured: the reference's
// PointStyle as its disc route reads it (modules/render/include/lrt/render/
// Points.h:21-33; shaders/lrt/points/point_frame.slang:7-44).
struct PointParams {
    uint32_t sizeMode;         // 0: size in world units, 1: in pixels
    float size;                // diameter
    uint32_t constantColour;   // 1: colour below for every point; 0: the cloud's own
    float colour[3];
};

// One camera of a batched frame as the duplication and the blend see it. A
// batch's views share one slot space (view v's splats at v * splatsPerView)
// and one tile space (view v's tiles from tileBase); pixelBase places the
// view's per-pixel blend state in the batch's arrays.
struct ViewInfo {
    uint32_t width, height, tilesX, tilesY;
    uint32_t tileBase;     // first tile of this view in the batch
    uint32_t params;       // index of its (first instance's) FrameParams
    uint64_t pixelBase;
    float* colour;         // float4 a pixel, the caller's
    float* depth;          // may be null
    // The depth keys (bits(z) >> 7) the batch's sort key holds: a drawn slot
    // outside [depthLo, depthHi] is an error the duplication counts.
    uint32_t depthLo, depthHi;
};

// What the projection leaves for the blend, one record per slot: the
// reference's ProjRecord (frame.slang:85-89), three float4 = 48 bytes.
//   screen = centre.xy, view z, cutoff power
//   conic  = a, b, c, alpha
//   colour = rgb (encoded), 0
struct ProjRecord {
    float screen[4];
    float conic[4];
    float colour[4];
};

// ---- depth-key ranges ------------------------------------------------------
// The depth keys (bits(view z) >> 7) of a frame's drawn slots, folded per
// projecting workgroup of kProjectBlock slots: words 2e and 2e + 1 of
// `entries` are the least and greatest over workgroup e's drawn slots
// (0xFFFFFFFF and 0 for none), e = first + the workgroup's index. Every
// workgroup of a launch writes its entry, so a frame's launches fill entries
// [0, their workgroups) and the scan reduces them to the frame's range. A null
// `entries` writes nothing.
constexpr uint32_t kProjectBlock = 256;
struct DepthBlocks {
    uint32_t* entries;
    uint32_t first;
};
inline uint32_t projectBlocks(uint32_t count) { return (count + kProjectBlock - 1) / kProjectBlock; }

}   // namespace lrt
