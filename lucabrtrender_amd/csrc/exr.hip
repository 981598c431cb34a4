// OpenEXR writer (host code): see exr.h. The file is the format's
// single-part scanline layout: magic and version, the header attributes
// (sorted by name), the chunk offset table, then one chunk per 16 lines. A ZIP
// chunk holds the lines' bytes split into even and odd bytes, delta-coded
// (each byte minus its predecessor, plus 128), deflated; a chunk that does not
// shrink is stored raw, as the format allows.
#include <zlib.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <vector>

#include "exr.h"
#include "packing.h"

namespace lrt::io {
namespace {

constexpr uint32_t kLinesPerChunk = 16;   // ZIP_COMPRESSION
enum PixelType : int32_t { kUint = 0, kHalf = 1, kFloat = 2 };

struct Writer {
    std::vector<uint8_t> bytes;
    void raw(const void* p, size_t n) {
        const auto* b = static_cast<const uint8_t*>(p);
        bytes.insert(bytes.end(), b, b + n);
    }
    template <class T>
    void put(T v) {
        raw(&v, sizeof v);
    }
    void text(const char* s) { raw(s, std::strlen(s) + 1); }
    void attribute(const char* name, const char* type, const std::vector<uint8_t>& value) {
        text(name);
        text(type);
        put<int32_t>(static_cast<int32_t>(value.size()));
        raw(value.data(), value.size());
    }
};

template <class T>
void append(std::vector<uint8_t>& v, T x) {
    const auto* b = reinterpret_cast<const uint8_t*>(&x);
    v.insert(v.end(), b, b + sizeof x);
}

struct Channel {
    const char* name;
    int32_t type;
    const float* data;   // engine layout: row-major, bottom row first
    uint32_t stride;     // floats between pixels
};

// ZIP's byte reordering and delta predictor, then deflate. Returns the raw
// bytes when deflating does not make them smaller.
std::vector<uint8_t> zipChunk(const std::vector<uint8_t>& raw) {
    const size_t n = raw.size();
    std::vector<uint8_t> t(n);
    size_t a = 0, b = (n + 1) / 2;
    for (size_t i = 0; i < n; ++i) t[(i & 1u) == 0 ? a++ : b++] = raw[i];
    for (size_t i = n; i-- > 1;) t[i] = static_cast<uint8_t>(t[i] - t[i - 1] + 128);
    uLongf size = compressBound(static_cast<uLong>(n));
    std::vector<uint8_t> out(size);
    if (compress(out.data(), &size, t.data(), static_cast<uLong>(n)) != Z_OK || size >= n) return raw;
    out.resize(size);
    return out;
}

}   // namespace

Status writeExr(const std::string& path, uint32_t width, uint32_t height, const float* rgba, const float* depth,
                bool half) {
    if (width == 0 || height == 0 || rgba == nullptr) return Status::error(ErrorCode::InvalidArgument, "no pixels to write");
    if (width > 0x7FFFFFFFu || height > 0x7FFFFFFFu) return Status::error(ErrorCode::InvalidArgument, "image too large");
    // Channels in name order: A B G R (Z).
    const int32_t colourType = half ? kHalf : kFloat;
    std::vector<Channel> channels{{"A", colourType, rgba + 3, 4}, {"B", colourType, rgba + 2, 4},
                                  {"G", colourType, rgba + 1, 4}, {"R", colourType, rgba + 0, 4}};
    if (depth != nullptr) channels.push_back({"Z", kFloat, depth, 1});

    Writer w;
    const uint32_t magic = 20000630u;
    w.put(magic);
    w.put<uint32_t>(2u);   // version 2, single-part scanline
    std::vector<uint8_t> chlist;
    for (const Channel& c : channels) {
        chlist.insert(chlist.end(), c.name, c.name + std::strlen(c.name) + 1);
        append<int32_t>(chlist, c.type);
        append<uint8_t>(chlist, 0);   // pLinear
        append<uint8_t>(chlist, 0);
        append<uint8_t>(chlist, 0);
        append<uint8_t>(chlist, 0);   // reserved
        append<int32_t>(chlist, 1);   // x sampling
        append<int32_t>(chlist, 1);   // y sampling
    }
    chlist.push_back(0);
    std::vector<uint8_t> box;
    append<int32_t>(box, 0);
    append<int32_t>(box, 0);
    append<int32_t>(box, static_cast<int32_t>(width) - 1);
    append<int32_t>(box, static_cast<int32_t>(height) - 1);
    std::vector<uint8_t> one, centre;
    append<float>(one, 1.0f);
    append<float>(centre, 0.0f);
    append<float>(centre, 0.0f);
    w.attribute("channels", "chlist", chlist);
    w.attribute("compression", "compression", {3});   // ZIP_COMPRESSION
    w.attribute("dataWindow", "box2i", box);
    w.attribute("displayWindow", "box2i", box);
    w.attribute("lineOrder", "lineOrder", {0});       // INCREASING_Y
    w.attribute("pixelAspectRatio", "float", one);
    w.attribute("screenWindowCenter", "v2f", centre);
    w.attribute("screenWindowWidth", "float", one);
    w.bytes.push_back(0);   // end of header

    const uint32_t chunks = (height + kLinesPerChunk - 1) / kLinesPerChunk;
    const size_t tableAt = w.bytes.size();
    w.bytes.resize(tableAt + size_t{chunks} * 8);
    std::vector<uint8_t> raw;
    for (uint32_t c = 0; c < chunks; ++c) {
        const uint32_t y0 = c * kLinesPerChunk, y1 = std::min(height, y0 + kLinesPerChunk);
        raw.clear();
        for (uint32_t y = y0; y < y1; ++y) {
            const size_t row = size_t{height - 1 - y} * width;   // the file's line y is the engine's row h-1-y
            for (const Channel& ch : channels) {
                for (uint32_t x = 0; x < width; ++x) {
                    const float v = ch.data[(row + x) * ch.stride];
                    if (ch.type == kHalf) {
                        append<uint16_t>(raw, pack::floatToHalf(v));
                    } else {
                        append<float>(raw, v);
                    }
                }
            }
        }
        const std::vector<uint8_t> data = zipChunk(raw);
        const uint64_t offset = w.bytes.size();
        std::memcpy(w.bytes.data() + tableAt + size_t{c} * 8, &offset, 8);
        w.put<int32_t>(static_cast<int32_t>(y0));
        w.put<int32_t>(static_cast<int32_t>(data.size()));
        w.raw(data.data(), data.size());
    }
    FILE* f = std::fopen(path.c_str(), "wb");
    if (f == nullptr) return Status::error(ErrorCode::IoFailure, "cannot write '" + path + "'");
    const bool ok = std::fwrite(w.bytes.data(), 1, w.bytes.size(), f) == w.bytes.size();
    if (std::fclose(f) != 0 || !ok) return Status::error(ErrorCode::IoFailure, "cannot write '" + path + "'");
    return {};
}

}   // namespace lrt::io
