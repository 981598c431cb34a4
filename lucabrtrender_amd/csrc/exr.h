// OpenEXR output: the reference's io::writeExr (modules/io/src/Exr.cpp:130-157)
// without its tinyexr dependency. One scanline image, ZIP compression (16
// lines a chunk, deflate via zlib), channels in name order as the format keeps
// them, rows written top first from the engine's bottom-first images:
// R, G, B, A as half (or float) and Z as float when depth is given.
#pragma once

#include <cstdint>
#include <string>

#include "status.h"

namespace lrt::io {

// rgba: width * height * 4 floats, depth: width * height floats or null; host
// memory, bottom row first.
Status writeExr(const std::string& path, uint32_t width, uint32_t height, const float* rgba, const float* depth,
                bool half);

}   // namespace lrt::io
