// Host splat readers: see readers.h. Files are memory-mapped and their
// records copied (all-float32 PLY, as the reference copies the file's own
// bytes, SplatReaders.cpp:180-186) or converted value by value, in parallel
// over blocks of 64k records.
#include "readers.h"

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cctype>
#include <cstring>
#include <sstream>
#include <thread>

#include <zlib.h>

namespace lrt::io {
namespace {

constexpr uint64_t kHeaderProbe = 1u << 16;
constexpr uint64_t kGrain = 1u << 16;   // records a parallel block

// A read-only mapping of a whole file.
class MappedFile {
public:
    ~MappedFile() {
        if (data_ != nullptr) munmap(const_cast<char*>(data_), size_);
        if (fd_ >= 0) close(fd_);
    }
    Status open(const std::string& path) {
        fd_ = ::open(path.c_str(), O_RDONLY);
        if (fd_ < 0) return Status::error(ErrorCode::IoFailure, "'" + path + "': cannot be opened");
        struct stat st {};
        if (fstat(fd_, &st) != 0) return Status::error(ErrorCode::IoFailure, "'" + path + "': cannot be read");
        size_ = static_cast<uint64_t>(st.st_size);
        if (size_ == 0) return {};
        void* p = mmap(nullptr, size_, PROT_READ, MAP_PRIVATE, fd_, 0);
        if (p == MAP_FAILED) return Status::error(ErrorCode::IoFailure, "'" + path + "': cannot be mapped");
        data_ = static_cast<const char*>(p);
        return {};
    }
    [[nodiscard]] const char* data() const { return data_; }
    [[nodiscard]] uint64_t size() const { return size_; }

private:
    int fd_ = -1;
    const char* data_ = nullptr;
    uint64_t size_ = 0;
};

uint32_t sizeOfType(const std::string& type) {
    if (type == "char" || type == "uchar" || type == "int8" || type == "uint8") return 1;
    if (type == "short" || type == "ushort" || type == "int16" || type == "uint16") return 2;
    if (type == "double" || type == "float64") return 8;
    return 4;
}

double readPlyValue(const char* record, const PlyProperty& p) {
    const char* at = record + p.offset;
    const auto get = [at](auto v) {
        std::memcpy(&v, at, sizeof v);
        return v;
    };
    const std::string& t = p.type;
    if (t == "float" || t == "float32") return static_cast<double>(get(float{}));
    if (t == "double" || t == "float64") return get(double{});
    if (t == "uchar" || t == "uint8") return get(uint8_t{});
    if (t == "char" || t == "int8") return get(int8_t{});
    if (t == "ushort" || t == "uint16") return get(uint16_t{});
    if (t == "short" || t == "int16") return get(int16_t{});
    if (t == "uint" || t == "uint32") return get(uint32_t{});
    return get(int32_t{});
}

// Runs body(first, last) over [0, count) in blocks of kGrain on the host's threads.
template <class Body>
void parallelBlocks(uint64_t count, Body body) {
    const uint64_t blocks = (count + kGrain - 1) / kGrain;
    const uint64_t threads = std::min<uint64_t>(blocks, std::max(1u, std::thread::hardware_concurrency()));
    if (threads <= 1) {
        body(0, count);
        return;
    }
    std::vector<std::thread> pool;
    for (uint64_t t = 0; t < threads; ++t) {
        pool.emplace_back([&, t] {
            for (uint64_t b = t; b < blocks; b += threads) body(b * kGrain, std::min(count, (b + 1) * kGrain));
        });
    }
    for (std::thread& th : pool) th.join();
}

std::string lowerExtension(const std::string& path) {
    const size_t slash = path.find_last_of('/');
    const size_t dot = path.find_last_of('.');
    if (dot == std::string::npos || (slash != std::string::npos && dot < slash)) return {};
    std::string ext = path.substr(dot);
    std::transform(ext.begin(), ext.end(), ext.begin(), [](unsigned char c) { return static_cast<char>(std::tolower(c)); });
    return ext;
}

std::string fileName(const std::string& path) {
    const size_t slash = path.find_last_of('/');
    return slash == std::string::npos ? path : path.substr(slash + 1);
}

}   // namespace

const PlyProperty* PlyElement::find(const std::string& property) const {
    for (const PlyProperty& p : properties) {
        if (p.name == property) return &p;
    }
    return nullptr;
}

const PlyElement* PlyHeader::element(const std::string& name) const {
    for (const PlyElement& e : elements) {
        if (e.name == name) return &e;
    }
    return nullptr;
}

Status parsePlyHeader(const char* text, uint64_t size, PlyHeader& out) {
    out = PlyHeader{};
    uint64_t at = 0;
    std::string line;
    const auto nextLine = [&]() {
        if (at >= size) return false;
        uint64_t end = at;
        while (end < size && text[end] != '\n') ++end;
        line.assign(text + at, end - at);
        at = end + 1;
        if (!line.empty() && line.back() == '\r') line.pop_back();
        return true;
    };
    if (!nextLine() || line.rfind("ply", 0) != 0) return Status::error(ErrorCode::InvalidArgument, "not a PLY file");
    bool ended = false;
    while (nextLine()) {
        std::istringstream words(line);
        std::string word;
        words >> word;
        if (word == "format") {
            words >> out.format;
            out.binaryLittleEndian = out.format == "binary_little_endian";
        } else if (word == "element") {
            PlyElement element;
            words >> element.name >> element.count;
            out.elements.push_back(std::move(element));
        } else if (word == "property") {
            if (out.elements.empty()) return Status::error(ErrorCode::InvalidArgument, "a PLY property before any element");
            PlyElement& element = out.elements.back();
            std::string type;
            words >> type;
            PlyProperty property;
            if (type == "list") {
                element.hasList = true;
                std::string countType, itemType;
                words >> countType >> itemType >> property.name;
                property.type = "list";
                element.properties.push_back(std::move(property));
                continue;
            }
            words >> property.name;
            property.type = type;
            property.offset = element.stride;
            property.size = sizeOfType(type);
            element.stride += property.size;
            element.properties.push_back(std::move(property));
        } else if (word == "end_header") {
            ended = true;
            break;
        }
    }
    if (!ended) return Status::error(ErrorCode::InvalidArgument, "PLY header has no end_header");
    if (out.format.empty()) return Status::error(ErrorCode::InvalidArgument, "PLY header has no format line");
    out.bytes = at;
    return {};
}

Status readSplats(const std::string& path, RawSplats& out) {
    const std::string ext = lowerExtension(path);
    if (ext == ".splat") return readDotSplat(path, out);
    if (ext == ".ply") return readSplatPly(path, out);
    if (ext == ".spz") return readSpz(path, out);
    if (ext == ".sog" || fileName(path) == "meta.json") {
        return Status::error(ErrorCode::Unsupported, "'" + path + "': this build reads no SOG image sets");
    }
    return Status::error(ErrorCode::Unsupported, "'" + path + "': no splat reader for '" + ext + "'");
}

Status readSplatPly(const std::string& path, RawSplats& out) {
    MappedFile file;
    Status s = file.open(path);
    if (!s.ok()) return s;
    PlyHeader header;
    s = parsePlyHeader(file.data(), std::min(file.size(), kHeaderProbe), header);
    if (!s.ok()) return Status::error(s.code, "'" + path + "': " + s.message);
    if (!header.binaryLittleEndian) {
        return Status::error(ErrorCode::Unsupported,
                             "'" + path + "': PLY format '" + header.format + "'; splat clouds are binary_little_endian");
    }
    const PlyElement* vertex = header.element("vertex");
    if (vertex == nullptr || vertex->count == 0) return Status::error(ErrorCode::InvalidArgument, "'" + path + "': no vertices");
    if (vertex->hasList) return Status::error(ErrorCode::Unsupported, "'" + path + "': a list in the vertex element");
    if (vertex->count > 0xFFFFFFFFull) return Status::error(ErrorCode::Unsupported, "'" + path + "': more than 2^32 splats");
    // The vertex element's offset: every element before it must be fixed-size.
    // Counts come from the header, so every product and sum is checked: a
    // header whose sizes wrap 64 bits or pass the file's end is refused.
    const Status ends = Status::error(ErrorCode::IoFailure, "'" + path + "': the file ends before its vertices do");
    uint64_t start = header.bytes;
    for (const PlyElement& e : header.elements) {
        if (e.name == "vertex") break;
        if (e.hasList) {
            return Status::error(ErrorCode::Unsupported,
                                 "'" + path + "': a list element ('" + e.name + "') before 'vertex' makes its offset unknowable");
        }
        uint64_t bytes = 0;
        if (__builtin_mul_overflow(uint64_t{e.stride}, e.count, &bytes) || __builtin_add_overflow(start, bytes, &start) ||
            start > file.size()) {
            return ends;
        }
    }
    uint64_t vertexBytes = 0, vertexEnd = 0;
    if (__builtin_mul_overflow(uint64_t{vertex->stride}, vertex->count, &vertexBytes) ||
        __builtin_add_overflow(start, vertexBytes, &vertexEnd) || vertexEnd > file.size()) {
        return ends;
    }
    static const char* const kRequired[] = {"x",     "y",     "z",     "opacity", "scale_0", "scale_1", "scale_2",
                                            "rot_0", "rot_1", "rot_2", "rot_3",   "f_dc_0",  "f_dc_1",  "f_dc_2"};
    for (const char* name : kRequired) {
        if (vertex->find(name) == nullptr) {
            return Status::error(ErrorCode::InvalidArgument,
                                 "'" + path + "': no '" + name + "' -- this does not look like a trained splat cloud");
        }
    }
    const bool allFloat32 = std::all_of(vertex->properties.begin(), vertex->properties.end(), [](const PlyProperty& p) {
        return p.type == "float" || p.type == "float32";
    });
    const auto floats = static_cast<uint32_t>(vertex->properties.size());
    const auto indexOf = [&](const std::string& name) {
        return static_cast<uint32_t>(vertex->find(name) - vertex->properties.data());
    };
    RawSplats raw;
    raw.source = path;
    raw.count = static_cast<uint32_t>(vertex->count);
    SplatEncoding& enc = raw.encoding;
    enc.floatsPerRecord = floats;
    enc.x = indexOf("x");
    enc.y = indexOf("y");
    enc.z = indexOf("z");
    enc.opacity = indexOf("opacity");
    enc.scale0 = indexOf("scale_0");
    enc.scale1 = indexOf("scale_1");
    enc.scale2 = indexOf("scale_2");
    enc.rotW = indexOf("rot_0");
    enc.rotX = indexOf("rot_1");
    enc.rotY = indexOf("rot_2");
    enc.rotZ = indexOf("rot_3");
    enc.dc0 = indexOf("f_dc_0");
    enc.dc1 = indexOf("f_dc_1");
    enc.dc2 = indexOf("f_dc_2");
    // The harmonics: counted, contiguous, colour-outermost; the degree is the
    // largest whole one the count covers (SplatReaders.cpp:134-168).
    uint32_t restCount = 0;
    for (const PlyProperty& p : vertex->properties) {
        if (p.name.rfind("f_rest_", 0) == 0) ++restCount;
    }
    uint32_t perColour = 0;
    for (uint32_t basis : {15u, 8u, 3u}) {
        if (restCount >= basis * 3) {
            perColour = basis;
            break;
        }
    }
    if (perColour > 0) {
        enc.restBase = indexOf("f_rest_0");
        const uint32_t filePerColour = restCount / 3;
        for (uint32_t k = 0; k < perColour * 3; ++k) {
            const uint32_t colour = k / perColour, basis = k % perColour;
            const uint32_t want = colour * filePerColour + basis;
            const std::string name = "f_rest_" + std::to_string(want);
            if (vertex->find(name) == nullptr || indexOf(name) != enc.restBase + want) {
                return Status::error(ErrorCode::Unsupported, "'" + path + "': harmonics are not contiguous in the vertex element");
            }
        }
        enc.restPerColour = perColour;
        enc.filePerColour = filePerColour;
        if (restCount / 3 != perColour) {
            return Status::error(ErrorCode::Unsupported, "'" + path + "': " + std::to_string(restCount) +
                                                             " harmonic coefficients is not a whole degree");
        }
    }
    enc.restColourOuter = 1;
    raw.records.resize(uint64_t{raw.count} * floats);
    const char* const base = file.data() + start;
    const uint64_t stride = vertex->stride;
    float* const dst = raw.records.data();
    parallelBlocks(raw.count, [&](uint64_t first, uint64_t last) {
        if (allFloat32) {
            std::memcpy(dst + first * floats, base + first * stride, (last - first) * stride);
            return;
        }
        for (uint64_t i = first; i < last; ++i) {
            const char* record = base + i * stride;
            float* o = dst + i * floats;
            for (uint32_t k = 0; k < floats; ++k) o[k] = static_cast<float>(readPlyValue(record, vertex->properties[k]));
        }
    });
    out = std::move(raw);
    return {};
}

Status readDotSplat(const std::string& path, RawSplats& out) {
    MappedFile file;
    Status s = file.open(path);
    if (!s.ok()) return s;
    constexpr uint64_t kRecord = 32;
    if (file.size() % kRecord != 0 || file.size() == 0) {
        return Status::error(ErrorCode::InvalidArgument, "'" + path + "': " + std::to_string(file.size()) +
                                                             " bytes is not a whole number of 32-byte .splat records");
    }
    if (file.size() / kRecord > 0xFFFFFFFFull) return Status::error(ErrorCode::Unsupported, "'" + path + "': more than 2^32 splats");
    // pos f32 x3 | scale f32 x3 | rgba u8 x4 | rot u8 x4 (w x y z, (v - 128) / 128)
    RawSplats raw;
    raw.source = path;
    raw.count = static_cast<uint32_t>(file.size() / kRecord);
    SplatEncoding& enc = raw.encoding;
    enc.floatsPerRecord = 14;
    enc.x = 0; enc.y = 1; enc.z = 2;
    enc.scale0 = 3; enc.scale1 = 4; enc.scale2 = 5;
    enc.dc0 = 6; enc.dc1 = 7; enc.dc2 = 8; enc.opacity = 9;
    enc.rotW = 10; enc.rotX = 11; enc.rotY = 12; enc.rotZ = 13;
    enc.opacityMode = SplatEncoding::Opacity::Byte;
    enc.scaleMode = SplatEncoding::Scale::Linear;
    enc.colourMode = SplatEncoding::Colour::Byte;
    enc.rotationMode = SplatEncoding::Rotation::Byte;
    raw.records.resize(uint64_t{raw.count} * enc.floatsPerRecord);
    const auto* data = reinterpret_cast<const unsigned char*>(file.data());
    float* const dst = raw.records.data();
    parallelBlocks(raw.count, [&](uint64_t first, uint64_t last) {
        for (uint64_t i = first; i < last; ++i) {
            const unsigned char* record = data + i * kRecord;
            float* o = dst + i * 14;
            std::memcpy(o, record, 24);   // six float32, as stored
            for (int k = 0; k < 8; ++k) o[6 + k] = static_cast<float>(record[24 + k]);
        }
    });
    out = std::move(raw);
    return {};
}

// SPZ (Niantic's packed splats), the reference's readSpz
// (modules/io/src/SplatReaders.cpp:242-328) over its loadSpzPacked: versions 1
// to 3 are one gzip stream holding a 16-byte header (magic "NGSP", version,
// points, SH degree, fractional bits, flags, a spare byte) and then the
// attribute streams one after another: positions (three 24-bit fixed-point
// integers a point; version 1's float16 positions were never released and are
// refused), alphas (a byte), colours (3 bytes), log scales (3 bytes),
// rotations (3 bytes in version 2, a smallest-three u32 in version 3) and the
// rest coefficients (3 bytes a basis, rgb per basis). Version 4 starts with the
// magic itself and keeps each stream in zstd: refused, as this build has no
// zstd. The records carry the bytes as values (a 24-bit integer is exact in a
// float; the smallest-three word as its two 16-bit halves), at most degree 3
// kept, and the encoding says how the GPU decodes them.
Status readSpz(const std::string& path, RawSplats& out) {
    constexpr uint32_t kMagic = 0x5053474eu;   // "NGSP"
    MappedFile file;
    Status s = file.open(path);
    if (!s.ok()) return s;
    const auto* bytes = reinterpret_cast<const unsigned char*>(file.data());
    const std::string where = "'" + path + "': ";
    if (file.size() >= 4) {
        uint32_t first = 0;
        std::memcpy(&first, bytes, 4);
        if (first == kMagic) {
            return Status::error(ErrorCode::Unsupported, where + "SPZ version 4 streams are zstd, which this build does not read");
        }
    }
    // The whole gzip stream, inflated.
    std::vector<unsigned char> data;
    {
        z_stream z{};
        if (inflateInit2(&z, 16 + MAX_WBITS) != Z_OK) return Status::error(ErrorCode::IoFailure, where + "zlib");
        z.next_in = const_cast<Bytef*>(reinterpret_cast<const Bytef*>(bytes));
        z.avail_in = static_cast<uInt>(std::min<uint64_t>(file.size(), 0xFFFFFFFFull));
        std::vector<unsigned char> chunk(1u << 20);
        int res = Z_OK;
        while (res == Z_OK) {
            z.next_out = chunk.data();
            z.avail_out = static_cast<uInt>(chunk.size());
            res = inflate(&z, Z_NO_FLUSH);
            if (res != Z_OK && res != Z_STREAM_END) break;
            data.insert(data.end(), chunk.data(), chunk.data() + (chunk.size() - z.avail_out));
        }
        inflateEnd(&z);
        if (res != Z_STREAM_END) return Status::error(ErrorCode::IoFailure, where + "not a readable SPZ file (gzip)");
    }
    if (data.size() < 16) return Status::error(ErrorCode::IoFailure, where + "not a readable SPZ file");
    uint32_t magic = 0, version = 0, points = 0;
    std::memcpy(&magic, data.data(), 4);
    std::memcpy(&version, data.data() + 4, 4);
    std::memcpy(&points, data.data() + 8, 4);
    const uint32_t shDegree = data[12], fractionalBits = data[13];
    if (magic != kMagic) return Status::error(ErrorCode::IoFailure, where + "not a readable SPZ file (magic)");
    if (version < 1 || version > 3) {
        return Status::error(ErrorCode::IoFailure, where + "SPZ version " + std::to_string(version) + " is not supported");
    }
    if (version == 1) {
        return Status::error(ErrorCode::Unsupported, where + "SPZ float16 positions (a format never released)");
    }
    if (points == 0) return Status::error(ErrorCode::IoFailure, where + "not a readable SPZ file (no points)");
    if (shDegree > 4 || fractionalBits > 24) {
        return Status::error(ErrorCode::IoFailure, where + "not a readable SPZ file (header)");
    }
    const uint32_t shDim = shDegree >= 4 ? 24u : shDegree == 3 ? 15u : shDegree == 2 ? 8u : shDegree == 1 ? 3u : 0u;
    const bool smallestThree = version >= 3;
    const uint64_t n = points;
    const uint64_t rotBytes = smallestThree ? 4 : 3;
    const uint64_t need = 16 + n * (9 + 1 + 3 + 3 + rotBytes + uint64_t{shDim} * 3);
    if (data.size() < need) {
        return Status::error(ErrorCode::IoFailure,
                             where + "SPZ attribute streams shorter than " + std::to_string(n) + " points");
    }
    const unsigned char* pos = data.data() + 16;
    const unsigned char* alpha = pos + n * 9;
    const unsigned char* colour = alpha + n;
    const unsigned char* scale = colour + n * 3;
    const unsigned char* rot = scale + n * 3;
    const unsigned char* rest = rot + n * rotBytes;

    RawSplats raw;
    raw.source = path;
    raw.count = points;
    SplatEncoding& enc = raw.encoding;
    const uint32_t kept = std::min(shDim, 15u);   // the engine evaluates up to degree 3
    enc.floatsPerRecord = 14 + kept * 3;
    enc.x = 0; enc.y = 1; enc.z = 2;
    enc.opacity = 3;
    enc.scale0 = 4; enc.scale1 = 5; enc.scale2 = 6;
    enc.dc0 = 7; enc.dc1 = 8; enc.dc2 = 9;
    enc.rotX = 10; enc.rotY = 11; enc.rotZ = 12; enc.rotW = 13;
    enc.restBase = 14;
    enc.restPerColour = kept;
    enc.filePerColour = kept;
    enc.restColourOuter = 0;   // rgb per basis
    enc.opacityMode = SplatEncoding::Opacity::Byte;
    enc.scaleMode = SplatEncoding::Scale::SpzByte;
    enc.colourMode = SplatEncoding::Colour::SpzByte;
    enc.rotationMode = smallestThree ? SplatEncoding::Rotation::SmallestThree : SplatEncoding::Rotation::FirstThree;
    enc.restMode = SplatEncoding::Rest::Byte;
    enc.fractionalBits = fractionalBits;
    enc.flipYZ = 1;
    raw.records.assign(n * enc.floatsPerRecord, 0.0f);
    float* const dst = raw.records.data();
    const uint32_t stride = enc.floatsPerRecord;
    parallelBlocks(n, [&](uint64_t first, uint64_t last) {
        for (uint64_t i = first; i < last; ++i) {
            float* o = dst + i * stride;
            const unsigned char* p = pos + i * 9;
            for (int k = 0; k < 3; ++k) {
                // A little-endian 24-bit two's-complement integer, kept unscaled.
                int32_t fixed = p[k * 3] | (p[k * 3 + 1] << 8) | (p[k * 3 + 2] << 16);
                if ((fixed & 0x800000) != 0) fixed |= static_cast<int32_t>(0xFF000000u);
                o[k] = static_cast<float>(fixed);
            }
            o[3] = alpha[i];
            for (int k = 0; k < 3; ++k) {
                o[4 + k] = scale[i * 3 + k];
                o[7 + k] = colour[i * 3 + k];
            }
            const unsigned char* r = rot + i * rotBytes;
            if (smallestThree) {
                o[10] = static_cast<float>(r[0] | (r[1] << 8));
                o[11] = static_cast<float>(r[2] | (r[3] << 8));
            } else {
                o[10] = r[0];
                o[11] = r[1];
                o[12] = r[2];
            }
            // The first `kept` bases; the file's own count says where the next point starts.
            const unsigned char* h = rest + i * uint64_t{shDim} * 3;
            for (uint32_t k = 0; k < kept * 3; ++k) o[14 + k] = h[k];
        }
    });
    out = std::move(raw);
    return {};
}

}   // namespace lrt::io
