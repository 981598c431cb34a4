"""`lrt render` and `lrt bench` on the GPU (CmdRender.cpp:63-294): the EXR the
command writes for a reference fixture is, pixel for pixel, the frame the
Python interface draws from the same file and the same default view (the
clouds' box centre, from 1.2 x its largest side in front, focal 35 mm, near
0.01), rounded to half as the EXR stores colour; and `bench --stages` prints
the reference's frame and summary lines."""
import os
import re
import subprocess

import numpy as np
import pytest
import torch

from lucabrtrender_amd import raster
from test_exr import read_exr

pytestmark = pytest.mark.gpu

LRT = os.path.join(os.path.dirname(__file__), "..", "lucabrtrender_amd", "lrt")
SH3 = os.path.join(os.path.dirname(__file__), "golden", "splats", "sh3.ply")


def _run(*args):
    print("lrt", *args, flush=True)
    p = subprocess.run([LRT, *args], capture_output=True, text=True, timeout=120)
    print(p.stdout, p.stderr, flush=True)
    return p.returncode, p.stdout


def _python_frame(width, height, degree=3):
    rast = raster.TileRasterizer()
    cloud, info = rast.decode_cloud(raster.read_splats(SH3), degree)
    lo, hi = np.array(info.bounds_min, np.float64), np.array(info.bounds_max, np.float64)
    target = (lo + hi) * 0.5
    eye = target + np.array([0.0, 0.0, max(float(np.max(hi - lo)), 1e-3) * 1.2])
    cam = raster.Camera.looking_at(tuple(eye), tuple(target))
    cam.lens = raster.Lens(focal=35.0, near_z=0.01)
    t = raster.RenderTargets(width, height)
    stats = rast.render(cam, [raster.SplatInstance(cloud)], raster.RenderSettings(width=width, height=height,
                                                                                  max_sh_degree=degree), t)
    torch.cuda.synchronize()
    return t.colour.cpu().numpy(), t.depth.cpu().numpy(), stats


@pytest.mark.parametrize("degree", [3, 1])
def test_render_writes_the_frame_the_interface_draws(tmp_path, degree):
    out = str(tmp_path / "out.exr")
    code, text = _run("render", "--splats", SH3, "--size", "160x120", "--degree", str(degree), "-o", out)
    assert code == 0 and f"wrote {out}" in text
    colour, depth, stats = _python_frame(160, 120, degree)
    assert re.search(rf"{stats.splats} splats \({stats.visible} visible\), {stats.pairs} pairs, 160x120: best", text)
    _, channels, planes, _ = read_exr(out)
    assert [n for n, _ in channels] == ["A", "B", "G", "R", "Z"]
    top_first = colour[::-1]
    for k, name in enumerate("RGBA"):
        np.testing.assert_array_equal(planes[name], top_first[..., k].astype(np.float16).view(np.uint16))
    np.testing.assert_array_equal(planes["Z"], depth[::-1].view(np.uint32))


def test_bench_prints_frames_and_stages():
    code, text = _run("bench", "--splats", SH3, "--size", "320x240", "--repeat", "4", "--stages")
    assert code == 0
    assert len(re.findall(r"^frame \d: [0-9.]+ ms$", text, re.M)) == 4
    assert re.search(r"64 splats \(\d+ visible\), \d+ pairs, 320x240: best [0-9.]+ ms \(project [0-9.]+, depth sort "
                     r"[0-9.]+, counts [0-9.]+, emit [0-9.]+, tile sort\+ranges [0-9.]+, blend [0-9.]+\)", text)


def test_the_reference_technique_writes_the_fp64_brute_force(tmp_path):
    out = str(tmp_path / "ref.exr")
    code, text = _run("render", "--splats", SH3, "--size", "96x64", "--technique", "reference", "-o", out)
    assert code == 0 and "reference, 96x64: best" in text
    colour, _, _ = _python_frame(96, 64)
    _, _, planes, _ = read_exr(out)
    r = planes["R"].view(np.float16).astype(np.float32)[::-1]
    assert np.max(np.abs(r - colour[..., 0])) <= 1e-2   # half rounding plus the raster's tolerance


def test_lod_frames_cut_every_frame_and_say_so(tmp_path):
    """CmdRender.cpp:279-417: built once, cut every frame; --lod with a threshold
    every cell meets draws the coarsest level alone, and threshold 0.001 the
    cloud as the Python cut draws it."""
    code, text = _run("bench", "--splats", SH3, "--size", "160x120", "--repeat", "3", "--lod", "1000000")
    assert code == 0 and re.search(r"^levels of detail built in [0-9.]+ ms$", text, re.M)
    assert len(re.findall(r"^frame \d: [0-9.]+ ms \(cut [0-9.]+ \+ render [0-9.]+\)$", text, re.M)) == 3
    m = re.search(r"^cut: 0 splats \+ (\d+) merged of 64, [0-9.]+ ms; (\d+) splats \(\d+ visible\)", text, re.M)
    assert m and m.group(1) == m.group(2) and int(m.group(1)) > 0
    out = str(tmp_path / "lod.exr")
    code, text = _run("render", "--splats", SH3, "--size", "160x120", "--lod", "0.001", "-o", out)
    assert code == 0 and "cut: 64 splats + 0 merged of 64" in text and f"wrote {out}" in text
    rast = raster.TileRasterizer()
    cloud, info = rast.decode_cloud(raster.read_splats(SH3), 3)
    lod = rast.lod_build(cloud)
    lo, hi = np.array(info.bounds_min, np.float64), np.array(info.bounds_max, np.float64)
    target = (lo + hi) * 0.5
    eye = target + np.array([0.0, 0.0, max(float(np.max(hi - lo)), 1e-3) * 1.2])
    cam = raster.Camera.looking_at(tuple(eye), tuple(target))
    cam.lens = raster.Lens(focal=35.0, near_z=0.01)
    s = raster.RenderSettings(width=160, height=120)
    drawn, _ = rast.lod_select(cam, [raster.LodInstance(lod)], 0.001, s)
    t = raster.RenderTargets(160, 120)
    rast.render(cam, drawn, s, t)
    torch.cuda.synchronize()
    _, _, planes, _ = read_exr(out)
    want = t.colour.cpu().numpy()[::-1]
    for k, name in enumerate("RGBA"):
        np.testing.assert_array_equal(planes[name].view(np.float16), want[..., k].astype(np.float16))
