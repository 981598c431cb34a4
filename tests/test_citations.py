"""Every reference citation in the boundary headers (include/*.h) points at the
symbol it names.

Each `file:first-last` citation in include/*.h must appear in CITED below with
the reference file it means and a text that must stand on one of its first
three lines (the declaration it cites). A citation added to a header without an
entry here fails, so does one whose lines no longer hold the symbol. The
reference sources are read as text; where they are not present the test is
skipped (they are not shipped with this repository).
"""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REFERENCE = "/root/reference"
HEADERS = [os.path.join(ROOT, "include", n) for n in sorted(os.listdir(os.path.join(ROOT, "include")))
           if n.endswith(".h")]
CITE = re.compile(r"([A-Za-z_][A-Za-z0-9_/.\-]*\.(?:h|cpp|slang)):(\d+)(?:-(\d+))?")

R = "modules/render/include/lrt/render/"
LOD = "modules/lod/include/lrt/lod/Lod.h"
# citation as written -> (reference file, text on one of its first three lines)
CITED = {
    "modules/core/include/lrt/core/Error.h:14-26": ("modules/core/include/lrt/core/Error.h", "enum class ErrorCode"),
    "modules/render/include/lrt/render/Camera.h:23-44": (R + "Camera.h", "struct Lens"),
    "Camera.h:46-53": (R + "Camera.h", "struct Camera"),
    "frame.slang:31": ("shaders/lrt/splat/frame.slang", "kTile = 16"),
    "Camera.h:56-80": (R + "Camera.h", "struct Projection"),
    "modules/render/src/Camera.cpp:9-13": ("modules/render/src/Camera.cpp", "Camera::lookingAt"),
    "Transform.h:376-400": ("modules/aofx/sdk/include/aofx/Transform.h", "lookAt("),
    "modules/render/src/Camera.cpp:20-48": ("modules/render/src/Camera.cpp", "projectionFor("),
    "modules/scene/include/lrt/scene/GpuClouds.h:41-76": ("modules/scene/include/lrt/scene/GpuClouds.h",
                                                          "struct GpuSplats"),
    "GpuClouds.h:41-76": ("modules/scene/include/lrt/scene/GpuClouds.h", "struct GpuSplats"),
    "shaders/lrt/common/packing.slang:5-11": ("shaders/lrt/common/packing.slang", "positions  float4"),
    "packing.slang:5-11": ("shaders/lrt/common/packing.slang", "positions  float4"),
    "modules/render/include/lrt/render/TileRasterizer.h:29-50": (R + "TileRasterizer.h", "struct SplatEdit"),
    "TileRasterizer.h:52-67": (R + "TileRasterizer.h", "struct SplatInstance"),
    "TileRasterizer.h:105-118": (R + "TileRasterizer.h", "struct RenderSettings"),
    "TileRasterizer.h:129-134": (R + "TileRasterizer.h", "struct RenderTargets"),
    "TileRasterizer.h:120-126": (R + "TileRasterizer.h", "struct FrameStats"),
    "TileRasterizer.h:138": (R + "TileRasterizer.h", "static Result<TileRasterizer> create("),
    "TileRasterizer.cpp:59-103": ("modules/render/src/TileRasterizer.cpp", "TileRasterizer::create("),
    "TileRasterizer.h:143-155": (R + "TileRasterizer.h", "Result<FrameStats> render("),
    "TileRasterizer.cpp:188-484": ("modules/render/src/TileRasterizer.cpp", "TileRasterizer::render("),
    "modules/scene/include/lrt/scene/GpuClouds.h:78-85": ("modules/scene/include/lrt/scene/GpuClouds.h",
                                                          "struct GpuPoints"),
    "modules/render/include/lrt/render/Points.h:21-33": (R + "Points.h", "struct PointStyle"),
    "Points.h:35-39": (R + "Points.h", "struct PointInstance"),
    "points_project.slang:33-86": ("shaders/lrt/splat/points_project.slang", "void pointsProject("),
    "splat_blend.slang:23-120": ("shaders/lrt/splat/splat_blend.slang", "void blendPixel("),
    "modules/lod/include/lrt/lod/Lod.h:40-143": (LOD, "struct LodLevel"),
    "Lod.h:78-87": (LOD, "struct LodBuildSettings"),
    "Lod.h:89-99": (LOD, "class LodBuilder"),
    "Lod.cpp:81-283": ("modules/lod/src/Lod.cpp", "LodBuilder::build("),
    "Lod.h:51-76": (LOD, "struct LodCloud"),
    "Lod.h:101-106": (LOD, "struct LodInstance"),
    "Lod.h:108-117": (LOD, "struct CutStats"),
    "Lod.h:131-134": (LOD, "select(const render::Projection&"),
    "Lod.cpp:350-547": ("modules/lod/src/Lod.cpp", "CutSelector::select("),
    "modules/render/src/ReferenceRenderer.cpp:54-160": ("modules/render/src/ReferenceRenderer.cpp",
                                                        "ReferenceRenderer::render("),
    "reference_project.slang:31-105": ("shaders/lrt/reference/reference_project.slang", "void referenceProject("),
    "reference_blend.slang:24-119": ("shaders/lrt/reference/reference_blend.slang", "void referenceBlend("),
    "modules/render/src/FrameParams.h:32-69": ("modules/render/src/FrameParams.h", "void setFrame("),
    "TileRasterizer.cpp:297-314": ("modules/render/src/TileRasterizer.cpp", "const Mat4 objectToView"),
    "modules/gpu/include/lrt/gpu/algo/PrefixSum.h:43-44": ("modules/gpu/include/lrt/gpu/algo/PrefixSum.h",
                                                           "Result<void> apply("),
    "modules/gpu/include/lrt/gpu/algo/RadixSort.h:56-57": ("modules/gpu/include/lrt/gpu/algo/RadixSort.h",
                                                           "Result<void> sort("),
    "modules/io/src/Exr.cpp:130-157": ("modules/io/src/Exr.cpp", "Result<void> writeExr("),
    "modules/io/include/lrt/io/RawSplats.h:15-77": ("modules/io/include/lrt/io/RawSplats.h", "struct SplatEncoding"),
    "RawSplats.h:42-58": ("modules/io/include/lrt/io/RawSplats.h", "enum class Opacity"),
    "modules/io/src/SplatReaders.cpp:32-195": ("modules/io/src/SplatReaders.cpp", "readSplats("),
    "modules/io/src/SplatReaders.cpp:32-328": ("modules/io/src/SplatReaders.cpp", "readSplats("),
    "modules/scene/src/GpuClouds.cpp:151-227": ("modules/scene/src/GpuClouds.cpp", "CloudLoader::decodeSlice("),
    "modules/scene/src/GpuClouds.cpp:199-227": ("modules/scene/src/GpuClouds.cpp", "CloudLoader::upload("),
}


def _citations():
    out = []
    for path in HEADERS:
        for number, line in enumerate(open(path), 1):
            for m in CITE.finditer(line):
                out.append((os.path.relpath(path, ROOT), number, m.group(0), int(m.group(2)),
                            int(m.group(3) or m.group(2))))
    return out


def test_every_header_citation_is_registered():
    missing = [(h, n, c) for h, n, c, _, _ in _citations() if c not in CITED]
    assert not missing, missing
    assert len(_citations()) >= 40


@pytest.mark.skipif(not os.path.isdir(REFERENCE), reason="the reference sources are not present here")
def test_every_header_citation_names_its_symbol():
    wrong = []
    for header, number, cite, first, last in _citations():
        path, text = CITED[cite]
        lines = open(os.path.join(REFERENCE, path), errors="replace").read().split("\n")
        if last > len(lines) or not any(text in lines[k - 1] for k in range(first, min(first + 3, len(lines) + 1))):
            wrong.append(f"{header}:{number}: {cite} does not hold {text!r}")
    assert not wrong, "\n".join(wrong)
