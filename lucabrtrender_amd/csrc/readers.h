// Splat files into raw records, on the host: the reference's io::readSplats
// (modules/io/src/SplatReaders.cpp:32-195), io::RawSplats and io::SplatEncoding
// (modules/io/include/lrt/io/RawSplats.h:15-80) and its PLY header parser
// (modules/io/src/PlyHeader.cpp). A file becomes one float a property a
// record, in the file's order, plus where each field sits and how it is
// encoded; the decode into packed splats runs on the GPU (load.hip).
//
// Read: binary little-endian PLY (3DGS properties), .splat (32-byte records)
// and SPZ up to version 3 (one gzip stream, through zlib). SPZ version 4 keeps
// its streams in zstd, which this build does not have, and .sog is a
// device-decoded image set: both are refused as Unsupported, as the reference
// refuses what its build lacks a library for.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "status.h"

namespace lrt::io {

struct SplatEncoding {
    static constexpr uint32_t kNoField = 0xFFFFFFFFu;
    uint32_t floatsPerRecord = 0;
    // Offsets, in floats, from the start of a record.
    uint32_t x = 0, y = 0, z = 0, opacity = 0;
    uint32_t scale0 = 0, scale1 = 0, scale2 = 0;
    uint32_t rotW = 0, rotX = 0, rotY = 0, rotZ = 0;
    uint32_t dc0 = 0, dc1 = 0, dc2 = 0;
    // First rest coefficient, basis functions per colour kept (0, 3, 8, 15)
    // and in the file (where green starts when colour-outer).
    uint32_t restBase = 0, restPerColour = 0, filePerColour = 0;
    uint32_t restColourOuter = 1;   // 1: rrr.. ggg.. bbb.. (3DGS PLY); 0: rgb per basis
    // Modes, numbered as RawSplats.h:42-58 numbers them.
    enum class Opacity : uint32_t { Logit = 0, Linear = 1, Byte = 2 };
    // SpzByte: log scale = byte / 16 - 10.
    enum class Scale : uint32_t { Log = 0, Linear = 1, SpzByte = 2 };
    // SpzByte: SH DC = (byte / 255 - 0.5) / 0.15.
    enum class Colour : uint32_t { ShDc = 0, Linear = 1, Byte = 2, SpzByte = 3 };
    // FirstThree: x y z bytes, v / 127.5 - 1, w from unit length (SPZ v2).
    // SmallestThree: a u32 as its low 16 bits at rotX and high 16 at rotY (SPZ v3).
    enum class Rotation : uint32_t { Float = 0, Byte = 1, FirstThree = 2, SmallestThree = 3 };
    enum class Rest : uint32_t { Float = 0, Byte = 1 };
    Opacity opacityMode = Opacity::Logit;
    Scale scaleMode = Scale::Log;
    Colour colourMode = Colour::ShDc;
    Rotation rotationMode = Rotation::Float;
    Rest restMode = Rest::Float;
    // Positions are multiplied by 2^-fractionalBits (SPZ's fixed point).
    uint32_t fractionalBits = 0;
    // The file is right-up-back (SPZ); the engine reads right-down-front: y and
    // z of positions, rotations and harmonics change sign.
    uint32_t flipYZ = 0;
};

struct RawSplats {
    std::string source;
    uint32_t count = 0;
    SplatEncoding encoding;
    std::vector<float> records;   // count * floatsPerRecord
};

// One PLY element's fixed-size properties (no lists in an element read here).
struct PlyProperty {
    std::string name, type;
    uint32_t offset = 0, size = 0;
};
struct PlyElement {
    std::string name;
    uint64_t count = 0;
    uint32_t stride = 0;
    bool hasList = false;
    std::vector<PlyProperty> properties;
    [[nodiscard]] const PlyProperty* find(const std::string& property) const;
};
struct PlyHeader {
    std::string format;
    bool binaryLittleEndian = false;
    uint64_t bytes = 0;   // the header's length: the body starts here
    std::vector<PlyElement> elements;
    [[nodiscard]] const PlyElement* element(const std::string& name) const;
};

// PlyHeader.cpp's parsePlyHeader: the header in the first bytes of a file.
Status parsePlyHeader(const char* text, uint64_t size, PlyHeader& out);

// readSplats by extension (.ply, .splat, .spz; .sog refused).
Status readSplats(const std::string& path, RawSplats& out);
Status readSplatPly(const std::string& path, RawSplats& out);
Status readDotSplat(const std::string& path, RawSplats& out);
Status readSpz(const std::string& path, RawSplats& out);

}   // namespace lrt::io
