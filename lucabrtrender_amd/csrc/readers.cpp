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
    if (ext == ".spz") {
        return Status::error(ErrorCode::Unsupported, "'" + path + "': this build reads no .spz (zstd was not found)");
    }
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

}   // namespace lrt::io
