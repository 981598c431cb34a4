"""Batches of views: every view of a batch draws, bit for bit, what the
single-view frame draws for it (the same records, the same per-tile pair order
behind a per-view tile offset, the same blend), under the grid blend, the
persistent blend and the captured-graph replay; the graph replays correctly
after its cameras change and between eager frames; the autotuner returns one
of its candidates."""
import numpy as np
import pytest
import torch

import fixtures as F
from lucabrtrender_amd import _abi, raster, scenes

pytestmark = pytest.mark.gpu

SIZES = [(256, 256), (250, 190), (512, 384), (96, 64), (640, 360)]


@pytest.fixture(scope="module")
def cloud():
    return F.random_cloud(3000, 7).decode(3).to_device()


def _cam(k, n):
    a = 2.0 * np.pi * k / n
    c = raster.Camera.looking_at((5.0 * np.sin(a), 1.0 + 0.3 * k, 5.0 * np.cos(a)), (0.0, 0.0, 0.0))
    c.lens = raster.Lens(focal=30.0 + 3.0 * k)
    return c


def _single(rast, cam, size, inst, settings):
    s = raster.RenderSettings(**{**settings.__dict__, "width": size[0], "height": size[1]})
    t = raster.RenderTargets(*size)
    stats = rast.render(cam, inst, s, t)
    torch.cuda.synchronize()
    return t.colour.cpu().numpy(), t.depth.cpu().numpy(), stats.pairs


def _batch(rast, cams, sizes, inst, settings):
    views = [raster.View(c, raster.RenderTargets(*sz)) for c, sz in zip(cams, sizes)]
    stats = rast.render_views(views, inst, settings)
    torch.cuda.synchronize()
    return [(v.targets.colour.cpu().numpy(), v.targets.depth.cpu().numpy()) for v in views], stats


@pytest.mark.parametrize("persistent,graph,depth_mode,tile", [(False, False, 0, 16), (True, False, 0, 16),
                                                              (False, True, 1, 16), (True, True, 0, 16),
                                                              (False, False, 0, 8), (True, True, 1, 8),
                                                              (True, False, 1, 32), (False, True, 0, 32)])
def test_each_view_is_the_single_view_frame(cloud, persistent, graph, depth_mode, tile):
    rast = raster.TileRasterizer()
    inst = [raster.SplatInstance(cloud)]
    # 8-pixel tiles with 64-record lists, the smallest of each.
    settings = raster.RenderSettings(persistent_blend=persistent, use_graph=graph, depth_mode=depth_mode,
                                     background=(0.1, 0.0, 0.2, 0.3), tile_size=tile,
                                     blend_batch=64 if tile == 8 else 256)
    cams = [_cam(k, len(SIZES)) for k in range(len(SIZES))]
    out, stats = _batch(rast, cams, SIZES, inst, settings)
    total = 0
    for cam, size, (colour, depth) in zip(cams, SIZES, out):
        c1, d1, pairs = _single(raster.TileRasterizer(), cam, size, inst, settings)
        total += pairs
        np.testing.assert_array_equal(colour.view(np.uint32), c1.view(np.uint32))
        np.testing.assert_array_equal(depth.view(np.uint32), d1.view(np.uint32))
    assert stats.pairs == total
    assert stats.splats == cloud.count * len(SIZES)


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("far", [60.0, float("inf")])
def test_views_with_their_own_near_and_far_planes(cloud, graph, far):
    """The batch's depth field spans the least near plane's key to the greatest
    far plane's (all 24 bits when a far plane is infinite): views with planes of
    their own (near planes that cut into the cloud, an orthographic view) each
    draw their single-view frame, bit for bit."""
    planes = [(0.5, 50.0, False), (0.01, 1e4 if far < 1e30 else far, False), (2.0, far, False), (4.5, 6.0, False),
              (0.1, 30.0, True)]
    cams = []
    for k, (near, far, ortho) in enumerate(planes):
        c = _cam(k, len(planes))
        c.lens = raster.Lens(focal=5.0 if ortho else 30.0 + 3.0 * k, orthographic=ortho, near_z=near, far_z=far)
        cams.append(c)
    inst = [raster.SplatInstance(cloud)]
    settings = raster.RenderSettings(use_graph=graph)
    rast = raster.TileRasterizer()
    out, stats = _batch(rast, cams, SIZES, inst, settings)
    total = 0
    for cam, size, (colour, depth) in zip(cams, SIZES, out):
        c1, d1, pairs = _single(raster.TileRasterizer(), cam, size, inst, settings)
        total += pairs
        np.testing.assert_array_equal(colour.view(np.uint32), c1.view(np.uint32))
        np.testing.assert_array_equal(depth.view(np.uint32), d1.view(np.uint32))
    assert stats.pairs == total and total > 0


def test_two_instances_and_the_batch_debug_buffers(cloud):
    rast = raster.TileRasterizer()
    o2w = np.diag([0.5, 0.5, 0.5, 1.0])
    o2w[:3, 3] = (0.8, 0.0, -0.5)
    inst = [raster.SplatInstance(cloud), raster.SplatInstance(cloud, tuple(o2w.reshape(-1)))]
    settings = raster.RenderSettings()
    cams = [_cam(k, 3) for k in range(3)]
    sizes = SIZES[:3]
    out, stats = _batch(rast, cams, sizes, inst, settings)
    keys = rast.debug(_abi.DEBUG_KEYS)
    ranges = rast.debug(_abi.DEBUG_RANGES).reshape(-1, 2)
    tiles = sum(((w + 15) // 16) * ((h + 15) // 16) for w, h in sizes)
    assert len(ranges) == tiles and len(keys) == stats.pairs
    tile_of = (keys >> np.uint64(32)).astype(np.int64)
    assert np.all(np.diff(tile_of) >= 0) and tile_of.max() < tiles
    for cam, size, (colour, _) in zip(cams, sizes, out):
        c1, _, _ = _single(raster.TileRasterizer(), cam, size, inst, settings)
        np.testing.assert_array_equal(colour.view(np.uint32), c1.view(np.uint32))


def test_graph_replays_follow_new_cameras_between_eager_frames(cloud):
    rast = raster.TileRasterizer()
    inst = [raster.SplatInstance(cloud)]
    graph = raster.RenderSettings(use_graph=True)
    eager = raster.RenderSettings()
    sizes = SIZES[:4]
    for replay in range(4):
        cams = [_cam(k + replay, 7) for k in range(len(sizes))]
        out, _ = _batch(rast, cams, sizes, inst, graph)
        # eager work on the same context between replays
        rast.render(_cam(3, 5), inst, raster.RenderSettings(width=128, height=96), raster.RenderTargets(128, 96))
        for cam, size, (colour, _) in zip(cams, sizes, out):
            c1, _, _ = _single(raster.TileRasterizer(), cam, size, inst, eager)
            np.testing.assert_array_equal(colour.view(np.uint32), c1.view(np.uint32))


def test_autotune_picks_one_of_its_candidates(cloud):
    rast = raster.TileRasterizer()
    inst = [raster.SplatInstance(cloud)]
    views = [raster.View(_cam(k, 4), raster.RenderTargets(*SIZES[k])) for k in range(4)]
    best, ms = rast.autotune_views(views, inst, raster.RenderSettings())
    assert best.blend_batch in (64, 128, 256) and best.tile_size in (8, 16, 32) and ms > 0.0
    out = rast.render_views(views, inst, best)
    assert out.pairs > 0


def test_clustered_scene_in_a_batch_of_eight_orbit_views():
    cfg = scenes.workload("scene-fwd-3M-1080p", count=200_000)
    dev = cfg.scene.to_device()
    inst = [raster.SplatInstance(dev)]
    settings = raster.RenderSettings(dilation=cfg.dilation, filter3d=cfg.filter3d, use_graph=True)
    cams = [scenes.orbit_camera(15.0, 45.0 * k) for k in range(8)]
    sizes = [(1920, 1080), (1280, 720), (960, 540), (512, 384), (256, 256), (1920, 1080), (640, 480), (800, 600)]
    rast = raster.TileRasterizer()
    out, _ = _batch(rast, cams, sizes, inst, settings)
    for cam, size, (colour, depth) in zip(cams, sizes, out):
        c1, d1, _ = _single(raster.TileRasterizer(), cam, size, inst, settings)
        np.testing.assert_array_equal(colour.view(np.uint32), c1.view(np.uint32))
        np.testing.assert_array_equal(depth.view(np.uint32), d1.view(np.uint32))
