// Host-side splat packing and the seeded synthetic scenes (synth.hip).
#pragma once

#include <cstdint>

namespace lrt::synth {

enum class SceneKind : uint32_t {
    Cube = 0,       // means uniform in [-1, 1]^3 + offset
    Clusters = 1,   // anisotropic clusters: centres uniform in a box, members normal about them
    Mixed = 2,      // Cube, with a fraction pulled toward the camera and drawn larger
};

struct SceneSpec {
    SceneKind kind = SceneKind::Cube;
    uint64_t seed = 0;
    uint32_t degree = 0;
    float offset[3] = {0.0f, 0.0f, -4.0f};
    float logScaleMean = -4.0f, logScaleSd = 0.5f;
    float opacityLogitMean = 0.0f, opacityLogitSd = 1.0f;
    float shSd = 0.3f;   // band l coefficients ~ N(0, shSd / l)
    uint32_t clusters = 256;
    float boxHalf = 10.0f;
    float clusterSdLo = 0.05f, clusterSdHi = 1.0f;
    float nearFraction = 0.2f, nearLogScaleMean = -1.0f, nearLogScaleSd = 0.5f;
};

// One splat before packing: linear opacity, log scales, (x, y, z, w) rotation
// (any length), base colour, rest coefficients rgb per basis.
struct SplatFields {
    float position[3];
    float opacity;
    float logScale[3];
    float rotation[4];
    float base[3];
    float rest[45];
};

void packSplat(const SplatFields& s, uint32_t restPerColour, uint32_t shWords, float* position, uint32_t* shape,
               uint32_t* sh);
void generate(const SceneSpec& spec, uint32_t count, float* positions, uint32_t* shape, uint32_t* sh);

}   // namespace lrt::synth
