// lrt: `lrt render` and `lrt bench` for splat files over the library's C ABI,
// the reference's apps/lrt/src/CmdRender.cpp:31-349 for the splat path: a
// cloud (or several), an eye, a target, a lens; the tile rasteriser (or the
// FP64 brute force with --technique reference); an EXR out, or timed frames.
// Options, defaults (1920x1080, focal 35 mm, near 0.01, degree 3, eye 1.2 x
// the clouds' extent in front of their centre) and the lines printed are the
// reference's. --lod draws each cloud through its levels of detail, cut
// every frame (CmdRender.cpp:275-398). Point files, the streamed .lrtc store
// and ray tracing are not this library's: their options are refused.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/lrt_hip.h"

namespace {

struct Options {
    std::vector<std::string> splats;
    std::string technique = "raster";
    std::string output = "out.exr";
    std::vector<double> eye, target, up{0.0, 1.0, 0.0}, rotateX, scale;
    double focal = 35.0, nearZ = 0.01;
    std::string size = "1920x1080";
    unsigned degree = 3;
    bool noAntialias = false, stages = false;
    int repeat = 1;
    double lod = 0.0;   // pixels a merged cell may span; 0 draws the splats
};

void usage(const char* command) {
    std::fprintf(stderr,
                 "usage: lrt %s --splats FILE [--splats FILE ...] [--technique raster|reference]\n"
                 "       [--eye X Y Z] [--target X Y Z] [--up X Y Z] [--rotate-x DEG] [--scale X Y Z]\n"
                 "       [--focal MM] [--size WxH] [--near Z] [--degree 0..3] [--no-antialias] [--lod PX]%s\n",
                 command, std::strcmp(command, "bench") == 0 ? " [--repeat N] [--stages]" : " [-o|--output EXR]");
}

bool numbers(int argc, char** argv, int& i, size_t count, std::vector<double>& out) {
    out.clear();
    for (size_t k = 0; k < count; ++k) {
        if (i + 1 >= argc) return false;
        char* end = nullptr;
        out.push_back(std::strtod(argv[++i], &end));
        if (end == argv[i] || *end != '\0') return false;
    }
    return true;
}

// Parses argv[2..]; returns false (and says why) on anything it does not know.
bool parse(int argc, char** argv, bool bench, Options& o) {
    for (int i = 2; i < argc; ++i) {
        const std::string a = argv[i];
        const auto next = [&](std::string& into) {
            if (i + 1 >= argc) return false;
            into = argv[++i];
            return true;
        };
        std::vector<double> v;
        std::string text;
        bool ok = true;
        if (a == "--splats") {
            ok = next(text);
            if (ok) o.splats.push_back(text);
        } else if (a == "--technique") {
            ok = next(o.technique);
        } else if (a == "--eye") {
            ok = numbers(argc, argv, i, 3, o.eye);
        } else if (a == "--target") {
            ok = numbers(argc, argv, i, 3, o.target);
        } else if (a == "--up") {
            ok = numbers(argc, argv, i, 3, o.up);
        } else if (a == "--rotate-x") {
            ok = numbers(argc, argv, i, 1, o.rotateX);
        } else if (a == "--scale") {
            ok = numbers(argc, argv, i, 3, o.scale);
        } else if (a == "--focal") {
            ok = numbers(argc, argv, i, 1, v) && (o.focal = v[0], true);
        } else if (a == "--near") {
            ok = numbers(argc, argv, i, 1, v) && (o.nearZ = v[0], true);
        } else if (a == "--size") {
            ok = next(o.size);
        } else if (a == "--degree") {
            ok = numbers(argc, argv, i, 1, v) && v[0] >= 0 && (o.degree = static_cast<unsigned>(v[0]), true);
        } else if (a == "--no-antialias") {
            o.noAntialias = true;
        } else if (!bench && (a == "-o" || a == "--output")) {
            ok = next(o.output);
        } else if (bench && a == "--repeat") {
            ok = numbers(argc, argv, i, 1, v) && (o.repeat = static_cast<int>(v[0]), true);
        } else if (bench && a == "--stages") {
            o.stages = true;
        } else if (a == "--lod") {
            std::vector<double> v;
            ok = numbers(argc, argv, i, 1, v) && v[0] >= 0.0;
            if (ok) o.lod = v[0];
        } else if (a == "--points" || a == "--stream-budget" || a == "--point-route" || a == "--point-size" ||
                   a == "--point-pixels" || a == "--edl" || a == "--surface") {
            std::fprintf(stderr, "%s: point files and streamed .lrtc clouds are not read by this build\n", a.c_str());
            return false;
        } else {
            std::fprintf(stderr, "unknown option '%s'\n", a.c_str());
            return false;
        }
        if (!ok) {
            std::fprintf(stderr, "%s: missing or malformed value\n", a.c_str());
            return false;
        }
    }
    return true;
}

bool hipOk(hipError_t e, const char* what) {
    if (e == hipSuccess) return true;
    std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
    return false;
}

// A device buffer freed on scope exit.
struct DeviceMemory {
    void* ptr = nullptr;
    ~DeviceMemory() {
        if (ptr != nullptr) (void)hipFree(ptr);
    }
    bool allocate(size_t bytes, const char* what) { return hipOk(hipMalloc(&ptr, std::max<size_t>(bytes, 4)), what); }
    template <class T>
    T* as() const {
        return static_cast<T*>(ptr);
    }
};

struct Cloud {
    DeviceMemory positions, shape, sh;
    lrt_splats splats{};
};

// Row-major 4x4 on column vectors: rotation about x by `degrees` (aofx rotationX).
void rotationX(double degrees, double m[16]) {
    const double r = degrees * M_PI / 180.0, c = std::cos(r), s = std::sin(r);
    const double out[16] = {1, 0, 0, 0, 0, c, -s, 0, 0, s, c, 0, 0, 0, 0, 1};
    std::memcpy(m, out, sizeof out);
}

void multiply(const double a[16], const double b[16], double out[16]) {
    double t[16];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) {
            double sum = 0.0;
            for (int k = 0; k < 4; ++k) sum += a[r * 4 + k] * b[k * 4 + c];
            t[r * 4 + c] = sum;
        }
    std::memcpy(out, t, sizeof t);
}

int run(const Options& o, bool bench) {
    uint32_t width = 0, height = 0;
    if (std::sscanf(o.size.c_str(), "%ux%u", &width, &height) != 2 || width == 0 || height == 0) {
        std::fprintf(stderr, "--size wants WIDTHxHEIGHT\n");
        return 1;
    }
    if (o.splats.empty()) {
        std::fprintf(stderr, "nothing to render: give --splats\n");
        return 1;
    }
    if (o.technique != "raster" && o.technique != "reference") {
        std::fprintf(stderr, "--technique wants raster or reference (this build has no ray tracer)\n");
        return 1;
    }
    if (o.lod > 0.0 && o.technique != "raster") {
        std::fprintf(stderr, "--lod draws splats through the rasteriser\n");
        return 1;
    }
    lrt_context* ctx = nullptr;
    if (lrt_context_create(0, nullptr, &ctx) != LRT_OK) {
        std::fprintf(stderr, "%s\n", lrt_last_error(nullptr));
        return 1;
    }
    struct Guard {
        lrt_context* c;
        ~Guard() { lrt_context_destroy(c); }
    } guard{ctx};

    std::vector<Cloud> clouds(o.splats.size());
    float lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    for (size_t k = 0; k < o.splats.size(); ++k) {
        lrt_raw_splats* raw = nullptr;
        if (lrt_splats_read(o.splats[k].c_str(), &raw) != LRT_OK) {
            std::fprintf(stderr, "%s\n", lrt_last_error(nullptr));
            return 1;
        }
        const uint32_t n = lrt_raw_splats_count(raw);
        const uint32_t words = lrt_cloud_sh_words(raw, o.degree);
        Cloud& c = clouds[k];
        lrt_cloud_info info{};
        const bool allocated = c.positions.allocate(size_t{n} * 16, "positions") &&
                               c.shape.allocate(size_t{n} * 16, "shape") &&
                               (words <= 1 || c.sh.allocate(size_t{n} * 4 * words, "harmonics"));
        const lrt_status st = allocated ? lrt_cloud_decode(ctx, raw, o.degree, c.positions.as<float>(),
                                                           c.shape.as<uint32_t>(), c.sh.as<uint32_t>(), &info)
                                        : LRT_OUT_OF_MEMORY;
        lrt_raw_splats_destroy(raw);
        if (st != LRT_OK) {
            if (allocated) std::fprintf(stderr, "%s\n", lrt_last_error(ctx));
            return 1;
        }
        c.splats = {c.positions.as<float>(), c.shape.as<uint32_t>(), words > 1 ? c.sh.as<uint32_t>() : nullptr,
                    nullptr, info.count, info.rest_per_colour, info.sh_words};
        for (int a = 0; a < 3; ++a) {
            lo[a] = k == 0 ? info.bounds_min[a] : std::min(lo[a], info.bounds_min[a]);
            hi[a] = k == 0 ? info.bounds_max[a] : std::max(hi[a], info.bounds_max[a]);
        }
    }

    double model[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    if (o.rotateX.size() == 1) rotationX(o.rotateX[0], model);
    if (o.scale.size() == 3) {
        const double s[16] = {o.scale[0], 0, 0, 0, 0, o.scale[1], 0, 0, 0, 0, o.scale[2], 0, 0, 0, 0, 1};
        multiply(model, s, model);
    }
    std::vector<lrt_splat_instance> instances(clouds.size());
    for (size_t k = 0; k < clouds.size(); ++k) {
        instances[k].splats = &clouds[k].splats;
        std::memcpy(instances[k].object_to_world, model, sizeof model);
        lrt_splat_edit_default(&instances[k].edit);
    }
    // The default view: the centre of every cloud's box through the model, from
    // 1.2 x the box's largest side in front of it.
    const double centre[3] = {(double{lo[0]} + hi[0]) * 0.5, (double{lo[1]} + hi[1]) * 0.5, (double{lo[2]} + hi[2]) * 0.5};
    const double extent = std::max({double{hi[0]} - lo[0], double{hi[1]} - lo[1], double{hi[2]} - lo[2]});
    double target[3], eye[3], up[3] = {o.up[0], o.up[1], o.up[2]};
    for (int a = 0; a < 3; ++a) {
        target[a] = o.target.size() == 3 ? o.target[a]
                                         : model[a * 4 + 0] * centre[0] + model[a * 4 + 1] * centre[1] +
                                               model[a * 4 + 2] * centre[2] + model[a * 4 + 3];
    }
    for (int a = 0; a < 3; ++a) eye[a] = o.eye.size() == 3 ? o.eye[a] : target[a] + (a == 2 ? std::max(extent, 1e-3) * 1.2 : 0.0);
    lrt_camera camera;
    lrt_camera_looking_at(eye, target, up, &camera);
    camera.lens.focal = o.focal;
    camera.lens.near_z = o.nearZ;
    lrt_projection proj;
    if (lrt_projection_for(&camera, width, height, &proj) != LRT_OK) {
        std::fprintf(stderr, "the camera gives no projection\n");
        return 1;
    }
    lrt_render_settings settings;
    lrt_render_settings_default(&settings);
    settings.width = width;
    settings.height = height;
    settings.antialias = o.noAntialias ? 0 : 1;
    settings.max_sh_degree = o.degree;
    settings.time_stages = o.stages ? 1 : 0;

    // Levels of detail: built once, then cut every frame.
    std::vector<lrt_lod_cloud*> lods;
    struct LodGuard {
        std::vector<lrt_lod_cloud*>& l;
        ~LodGuard() {
            for (lrt_lod_cloud* c : l) lrt_lod_destroy(c);
        }
    } lodGuard{lods};
    std::vector<lrt_lod_instance> lodInstances;
    if (o.lod > 0.0) {
        const auto start = std::chrono::steady_clock::now();
        for (const Cloud& c : clouds) {
            lrt_lod_cloud* built = nullptr;
            if (lrt_lod_build(ctx, &c.splats, nullptr, &built) != LRT_OK) {
                std::fprintf(stderr, "%s\n", lrt_last_error(ctx));
                return 1;
            }
            lods.push_back(built);
            lrt_lod_instance li{};
            li.cloud = built;
            std::memcpy(li.object_to_world, model, sizeof model);
            lrt_splat_edit_default(&li.edit);
            lodInstances.push_back(li);
        }
        std::printf("levels of detail built in %.1f ms\n",
                    std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - start).count());
    }

    const size_t pixels = size_t{width} * height;
    std::vector<float> colour(pixels * 4), depth(pixels);
    const int rounds = std::max(o.repeat, 1);
    if (o.technique == "raster") {
        DeviceMemory colourDev, depthDev;
        if (!colourDev.allocate(pixels * 16, "colour") || !depthDev.allocate(pixels * 4, "depth")) return 1;
        lrt_render_targets targets{width, height, colourDev.as<float>(), depthDev.as<float>()};
        lrt_frame_stats best{};
        best.total_ms = 1e30;
        double bestCutMs = 0.0;
        lrt_cut_stats bestCut{};
        std::vector<lrt_splat_instance> drawn(std::max(lodInstances.size(), size_t{1}));
        for (int round = 0; round < rounds; ++round) {
            const lrt_splat_instance* list = instances.data();
            uint32_t count = static_cast<uint32_t>(instances.size());
            double cutMs = 0.0;
            std::vector<lrt_cut_stats> cutStats(lodInstances.size());
            if (!lodInstances.empty()) {
                const auto start = std::chrono::steady_clock::now();
                if (lrt_lod_select(ctx, &proj, lodInstances.data(), static_cast<uint32_t>(lodInstances.size()),
                                   static_cast<float>(o.lod), drawn.data(), &count, cutStats.data()) != LRT_OK) {
                    std::fprintf(stderr, "%s\n", lrt_last_error(ctx));
                    return 1;
                }
                cutMs = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - start).count();
                list = drawn.data();
            }
            lrt_frame_stats stats{};
            if (lrt_raster_render(ctx, &proj, list, count, &settings, &targets, &stats) != LRT_OK) {
                std::fprintf(stderr, "%s\n", lrt_last_error(ctx));
                return 1;
            }
            if (stats.total_ms + cutMs < best.total_ms + bestCutMs) {
                best = stats;
                bestCutMs = cutMs;
                if (!cutStats.empty()) bestCut = cutStats.front();
            }
            if (bench && !lodInstances.empty()) {
                std::printf("frame %d: %.2f ms (cut %.2f + render %.2f)\n", round, stats.total_ms + cutMs, cutMs,
                            stats.total_ms);
            } else if (bench) {
                std::printf("frame %d: %.2f ms\n", round, stats.total_ms);
            }
        }
        if (!lodInstances.empty()) {
            std::printf("cut: %u splats + %u merged of %u, %.2f ms; ", bestCut.splats, bestCut.merged,
                        bestCut.available, bestCutMs);
        }
        std::printf("%u splats (%u visible), %u pairs, %ux%u: best %.2f ms", best.splats, best.visible, best.pairs,
                    width, height, best.total_ms);
        if (o.stages) {
            std::printf(" (project %.2f, depth sort %.2f, counts %.2f, emit %.2f, tile sort+ranges %.2f, blend %.2f)",
                        best.project_ms, best.depth_sort_ms, best.counts_ms, best.emit_ms, best.tile_sort_ms,
                        best.blend_ms);
        }
        std::printf("\n");
        if (!bench && (!hipOk(hipMemcpy(colour.data(), colourDev.ptr, pixels * 16, hipMemcpyDeviceToHost), "read colour") ||
                       !hipOk(hipMemcpy(depth.data(), depthDev.ptr, pixels * 4, hipMemcpyDeviceToHost), "read depth"))) {
            return 1;
        }
    } else {
        DeviceMemory colourDev, depthDev;
        if (!colourDev.allocate(pixels * 32, "colour") || !depthDev.allocate(pixels * 8, "depth")) return 1;
        double bestMs = 1e30;
        for (int round = 0; round < rounds; ++round) {
            const auto start = std::chrono::steady_clock::now();
            if (lrt_reference_render(ctx, &proj, instances.data(), static_cast<uint32_t>(instances.size()), &settings,
                                     colourDev.as<double>(), depthDev.as<double>(), 0) != LRT_OK) {
                std::fprintf(stderr, "%s\n", lrt_last_error(ctx));
                return 1;
            }
            const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - start).count();
            if (bench) std::printf("frame %d: %.2f ms\n", round, ms);
            bestMs = std::min(bestMs, ms);
        }
        std::printf("%s, %ux%u: best %.2f ms\n", o.technique.c_str(), width, height, bestMs);
        if (!bench) {
            std::vector<double> c64(pixels * 4), d64(pixels);
            if (!hipOk(hipMemcpy(c64.data(), colourDev.ptr, pixels * 32, hipMemcpyDeviceToHost), "read colour") ||
                !hipOk(hipMemcpy(d64.data(), depthDev.ptr, pixels * 8, hipMemcpyDeviceToHost), "read depth")) {
                return 1;
            }
            for (size_t p = 0; p < pixels * 4; ++p) colour[p] = static_cast<float>(c64[p]);
            for (size_t p = 0; p < pixels; ++p) depth[p] = static_cast<float>(d64[p]);
        }
    }
    if (!bench) {
        if (lrt_write_exr(o.output.c_str(), width, height, colour.data(), depth.data(), 1) != LRT_OK) {
            std::fprintf(stderr, "%s\n", lrt_last_error(nullptr));
            return 1;
        }
        std::printf("wrote %s\n", o.output.c_str());
    }
    return 0;
}

}   // namespace

int main(int argc, char** argv) {
    const std::string command = argc > 1 ? argv[1] : "";
    if (command != "render" && command != "bench") {
        std::fprintf(stderr, "lrt: Gaussian splats on the GPU\nusage: lrt render|bench --splats FILE ... (--help)\n");
        return argc > 1 && (command == "--help" || command == "-h") ? 0 : 1;
    }
    for (int i = 2; i < argc; ++i) {
        if (std::strcmp(argv[i], "--help") == 0 || std::strcmp(argv[i], "-h") == 0) {
            usage(command.c_str());
            return 0;
        }
    }
    const bool bench = command == "bench";
    Options o;
    if (bench) o.repeat = 20;
    if (!parse(argc, argv, bench, o)) {
        usage(command.c_str());
        return 1;
    }
    return run(o, bench);
}
