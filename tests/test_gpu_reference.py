"""The FP64 brute force on the GPU (lrt_reference_render), the error oracle of
whole frames.

It agrees with the C restatement of the reference's brute force in FP64
(or_reference_render_f64, pinned on the CPU by tests/test_oracle_raster.py) on
the reference's own scene and cameras (randomCloud(3000, 7) at 250x190,
tests/render/test_splat_render.cpp:70-121), in both orders; then it judges the
tile rasteriser's full-resolution frames: p99 <= 1e-3 and max <= 1e-2 absolute
per channel in linear FP32 (the figures are printed).
"""
import numpy as np
import pytest
import torch

import fixtures as F
import harness as H
import oracle_lib as O
from lucabrtrender_amd import raster, scenes

pytestmark = pytest.mark.gpu

W, HGT = 250, 190


@pytest.fixture(scope="module")
def rast():
    return raster.TileRasterizer()


@pytest.fixture(scope="module")
def cloud():
    return F.random_cloud(3000, 7).decode(3)


def _cam(eye, target, **lens):
    c = raster.Camera.looking_at(eye, target)
    c.lens = raster.Lens(**lens)
    return c


CAMERAS = [_cam((1.5, 1.0, 6.0), (0, 0, 0), focal=30.0), _cam((0.3, -0.2, 2.2), (0, 0, 0), focal=18.0),
           _cam((0.0, 3.0, 5.0), (0, 0, 0), orthographic=True, focal=5.0)]


@pytest.mark.parametrize("keyed", [False, True])
@pytest.mark.parametrize("cam_index", range(len(CAMERAS)))
def test_gpu_fp64_brute_force_is_the_c_one(rast, cloud, cam_index, keyed):
    cam = CAMERAS[cam_index]
    s = raster.RenderSettings(width=W, height=HGT, background=(0.05, 0.1, 0.0, 0.25))
    colour, depth = rast.reference_render(cam, [raster.SplatInstance(cloud.to_device())], s, keyed=keyed)
    frames, clouds, n = H.oracle_frames(cam, s, [(cloud, np.eye(4), raster.SplatEdit())])
    c_colour, c_depth = O.reference(frames, clouds, n, f64=True, keyed=keyed)
    d = np.abs(colour.cpu().numpy() - c_colour).reshape(-1)
    print(f"camera {cam_index} keyed {keyed}: GPU vs C FP64 brute force p99 {np.quantile(d, 0.99):.1e} "
          f"max {d.max():.1e}")
    assert np.quantile(d, 0.99) <= 1e-10 and d.max() <= 1e-2
    dd = np.abs(depth.cpu().numpy() - c_depth).reshape(-1)
    assert np.quantile(dd, 0.99) <= 1e-9


def test_rasteriser_frame_against_the_fp64_brute_force_at_1080p(rast):
    """Against the brute force in its own order, (view z, index), p99 <= 1e-3;
    against it in the rasteriser's order (24-bit depth key, index), which the
    reference's tile path shares, also max <= 1e-2. Splats whose depths tie
    under the 24-bit key but not exactly are blended in another order by the
    two, which is where the first comparison's largest differences sit."""
    cfg = scenes.workload("scene-fwd-3M-1080p", count=200_000)
    s = raster.RenderSettings(width=cfg.width, height=cfg.height, dilation=cfg.dilation, filter3d=cfg.filter3d)
    inst = [raster.SplatInstance(cfg.scene.to_device())]
    targets = raster.RenderTargets(cfg.width, cfg.height)
    rast.render(cfg.camera, inst, s, targets)
    ours = targets.colour.cpu().numpy().astype(np.float64)
    for keyed in (False, True):
        colour, _ = rast.reference_render(cfg.camera, inst, s, keyed=keyed)
        torch.cuda.synchronize()
        d = np.abs(ours - colour.cpu().numpy()).reshape(-1)
        p99, mx = float(np.quantile(d, 0.99)), float(d.max())
        print(f"clustered scene (200k splats) 1920x1080 vs FP64 brute force, keyed {keyed}: p99 {p99:.2e} max {mx:.2e} "
              f"over 1e-3: {np.mean(d > 1e-3):.2e}")
        assert p99 <= 1e-3
        if keyed:
            assert mx <= 1e-2


def test_pixel_subset_is_the_whole_frame_at_those_pixels(rast, cloud):
    """lrt_reference_render_pixels computes, at each listed pixel, exactly what
    the whole-frame brute force computes there (same arithmetic, FP64)."""
    s = raster.RenderSettings(width=W, height=HGT, background=(0.05, 0.1, 0.0, 0.25))
    inst = [raster.SplatInstance(cloud.to_device())]
    pix = np.random.default_rng(3).choice(W * HGT, 700, replace=False).astype(np.int32)
    pix[:3] = [0, W * HGT - 1, W * 77 + 13]
    for keyed in (False, True):
        full, full_d = rast.reference_render(CAMERAS[0], inst, s, keyed=keyed)
        sub, sub_d = rast.reference_render_pixels(CAMERAS[0], inst, s, torch.from_numpy(pix).cuda(), keyed=keyed)
        f = full.reshape(-1, 4).cpu().numpy()[pix]
        assert np.array_equal(sub.cpu().numpy(), f)
        assert np.array_equal(sub_d.cpu().numpy(), full_d.reshape(-1).cpu().numpy()[pix])


def test_headline_3m_frame_against_the_fp64_brute_force_on_a_pixel_subset(rast):
    """The benchmark's frame itself (3,000,000 clustered splats, 1920x1080, 3D
    filter, 2D filter 0.1) at 65,536 seeded pixels against the FP64 brute force
    over all 3M splats. In the rasteriser's order (24-bit depth key, slot):
    p99 <= 1e-3 and max <= 1e-2. In the brute force's own order (exact view z,
    index): p99 <= 1e-3 and at most 1 in 1,000 channel values off by more than
    1e-2. Its max is not held to 1e-2: splats whose depths tie under the 24-bit
    key, which the reference's tile path sorts by and whose pairs here are bit
    for bit the reference's, are blended in slot order there and in exact depth
    order in the brute force; the measured max (about 0.1) is printed."""
    cfg = scenes.workload("scene-fwd-3M-1080p")
    s = raster.RenderSettings(width=cfg.width, height=cfg.height, dilation=cfg.dilation, filter3d=cfg.filter3d)
    inst = [raster.SplatInstance(cfg.scene.to_device())]
    targets = raster.RenderTargets(cfg.width, cfg.height)
    rast.render(cfg.camera, inst, s, targets)
    pix = np.random.default_rng(11).choice(cfg.width * cfg.height, 65536, replace=False).astype(np.int32)
    ours = targets.colour.reshape(-1, 4).cpu().numpy().astype(np.float64)[pix]
    for keyed in (True, False):
        print(f"3M-splat frame, 65536 pixels, FP64 brute force keyed {keyed} ...", flush=True)
        colour, _ = rast.reference_render_pixels(cfg.camera, inst, s, torch.from_numpy(pix).cuda(), keyed=keyed)
        d = np.abs(ours - colour.cpu().numpy()).reshape(-1)
        p99, mx = float(np.quantile(d, 0.99)), float(d.max())
        print(f"  p99 {p99:.2e} max {mx:.2e} over 1e-3: {np.mean(d > 1e-3):.2e}")
        assert p99 <= 1e-3
        if keyed:
            assert mx <= 1e-2
        else:
            assert np.mean(d > 1e-2) <= 1e-3 and mx <= 1.0
