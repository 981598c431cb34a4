"""The HIP disc route and composite blend against the oracle on the same
inputs: disc records, tile counts, sorted pairs and ranges bit for bit; images
within the blend's tolerance of the oracle's restatement and of the brute force
(the reference's tests/render/test_points.cpp:75-99, p99 <= 2); splats over a
disc layer as the under-layer against splats and discs in one sort (:136-197,
p99 <= 2); the under-layer's cut where the first pair at or behind a pixel's
under depth is one its quadrant never draws; and the call's argument checks."""
import numpy as np
import pytest

import fixtures as F
import harness as H
import oracle_lib as O
from lucabrtrender_amd import raster
from lucabrtrender_amd._abi import LrtError

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rast():
    return raster.TileRasterizer()


def _bit_exact(gpu, ora):
    n = ora["touched"].size
    assert gpu["counts"][0] == ora["splats"]
    assert gpu["counts"][1] == ora["visible"]
    assert gpu["counts"][2] == ora["pairs"]
    np.testing.assert_array_equal(gpu["touched"][:n], ora["touched"])
    vis = ora["touched"] > 0
    np.testing.assert_array_equal(gpu["proj"][:n][vis].view(np.uint32), ora["proj"][vis].view(np.uint32))
    np.testing.assert_array_equal((gpu["keys"] >> np.uint64(32)).astype(np.uint32), ora["pair_tiles"])
    np.testing.assert_array_equal(gpu["values"], ora["pair_splats"])
    np.testing.assert_array_equal(gpu["ranges"], ora["ranges"])


@pytest.mark.parametrize("mode,size", [(0, 0.06), (1, 5.0)])
def test_discs_match_the_oracle_and_the_brute_force(rast, mode, size):
    print(f"discs mode {mode} size {size}", flush=True)
    cam = raster.Camera.looking_at((1.0, 2.0, 6.0), (0.0, 0.0, 0.0))
    s = raster.RenderSettings(width=230, height=170, linearise=False)
    pos, col = F.random_points(4000, 3)
    pts = [(pos, col, np.eye(4), raster.PointStyle(size_mode=mode, size=size))]
    gpu = H.run_gpu_composite(cam, s, [], pts, rast=rast)
    ora = H.run_oracle_composite(cam, s, [], pts)
    _bit_exact(gpu, ora)
    p99, mx, over2 = O.compare_images(gpu["colour"], ora["colour"])
    assert p99 == 0 and over2 <= 3, (p99, mx, over2)
    truth, _ = H.run_reference_points(cam, s, [], pts)
    p99, mx, over2 = O.compare_images(gpu["colour"], truth)
    assert p99 <= 2, (p99, mx, over2)


@pytest.mark.parametrize("tile", [16, 8, 32])
def test_constant_colour_placed_instance_and_splats_mixed(rast, tile):
    """Splats and discs in one sort, at the reference's 16-pixel tiles and at
    8 and 32 (the oracle drawing at the same size)."""
    print(f"mixed splats and discs, tile {tile}", flush=True)
    cam = raster.Camera.looking_at((1.5, 1.0, 6.0), (0.0, 0.0, 0.0))
    s = raster.RenderSettings(width=250, height=190, depth_mode=1, tile_size=tile)
    scene = F.random_cloud(3000, 7).decode(3)
    pos, col = F.random_points(3000, 11)
    o2w = np.eye(4)
    o2w[:3, 3] = (0.3, -0.2, 0.5)
    pts = [(pos, col, np.eye(4), raster.PointStyle(size=0.03)),
           (pos[:800].copy(), None, o2w, raster.PointStyle(size_mode=1, size=3.0, constant_colour=True,
                                                            colour=(0.9, 0.1, 0.4)))]
    spl = [(scene, np.eye(4), raster.SplatEdit())]
    gpu = H.run_gpu_composite(cam, s, spl, pts, rast=rast)
    ora = H.run_oracle_composite(cam, s, spl, pts)
    _bit_exact(gpu, ora)
    p99, mx, over2 = O.compare_images(gpu["colour"], ora["colour"])
    assert p99 <= 1 and mx <= 16, (p99, mx, over2)
    truth, _ = H.run_reference_points(cam, s, spl, pts)
    p99, mx, over2 = O.compare_images(gpu["colour"], truth)
    assert p99 <= 2, (p99, mx, over2)
    # threshold depth: the first contributor past half coverage, or the disc's
    close = np.isclose(gpu["depth"], ora["depth"], rtol=1e-4, atol=1e-5)
    assert close.mean() > 0.995


def test_splats_over_a_disc_layer_match_splats_and_discs_in_one_sort(rast):
    print("under-layer against one sort", flush=True)
    cam = raster.Camera.looking_at((0.5, 0.5, 7.0), (0.0, 0.0, 0.0))
    s = raster.RenderSettings(width=256, height=180)
    pos, col = F.random_points(8000, 5)
    fog = F.decode_encoded(F.fog_records(), F.FOG_ENCODING)
    pts = [(pos, col, np.eye(4), raster.PointStyle(size=0.05))]
    spl = [(fog, np.eye(4), raster.SplatEdit())]
    layer = H.run_gpu_composite(cam, s, [], pts, rast=rast)
    over = H.run_gpu_composite(cam, s, spl, [], under=(layer["colour"], layer["depth"]), rast=rast)
    mixed = H.run_gpu_composite(cam, s, spl, pts, rast=rast)
    p99, mx, over2 = O.compare_images(over["colour"], mixed["colour"])
    assert p99 <= 2, (p99, mx, over2)
    ora = H.run_oracle_composite(cam, s, spl, [], layer["colour"], layer["depth"])
    p99, mx, over2 = O.compare_images(over["colour"], ora["colour"])
    assert p99 <= 1 and mx <= 16, (p99, mx, over2)


def test_the_under_layer_is_composited_with_the_transmittance_left(rast):
    print("under-layer arithmetic", flush=True)
    cam = raster.Camera.looking_at((1.5, 1.0, 6.0), (0.0, 0.0, 0.0))
    s = raster.RenderSettings(width=96, height=80)
    scene = F.random_cloud(400, 7).decode(3)
    spl = [(scene, np.eye(4), raster.SplatEdit())]
    under_colour = np.empty((s.height, s.width, 4), np.float32)
    under_colour[...] = (0.2, 0.3, 0.4, 1.0)
    for depth in (1000.0, 0.5, 0.0):
        under_depth = np.full((s.height, s.width), depth, np.float32)
        gpu = H.run_gpu_composite(cam, s, spl, [], under=(under_colour, under_depth), rast=rast)
        ora = H.run_oracle_composite(cam, s, spl, [], under_colour, under_depth)
        if depth == 0.5:   # nearer than every splat: the layer alone, exactly
            np.testing.assert_array_equal(gpu["colour"], under_colour)
            np.testing.assert_array_equal(gpu["depth"], under_depth)
        np.testing.assert_allclose(gpu["colour"], ora["colour"], rtol=0, atol=2e-5)
        np.testing.assert_allclose(gpu["depth"], ora["depth"], rtol=5e-5, atol=1e-3)


def test_a_pair_the_quadrant_never_draws_still_cuts_at_the_under_depth(rast):
    """Tile (1, 1) holds a tiny splat and, a few ulps nearer with the same depth
    key, a wide one after it. With the under depth at the tiny one's depth,
    every pixel of that tile stops at the tiny one, including the three
    quadrants the tiny one never reaches and whose lists drop it."""
    print("cut by a dropped pair", flush=True)
    cam = raster.Camera.looking_at((0.0, 0.0, 10.0), (0.0, 0.0, 0.0))
    cam.lens = raster.Lens(orthographic=True, focal=5.0)
    s = raster.RenderSettings(width=64, height=64, linearise=False)
    b = F.CloudBuilder()
    b.add(-0.9, -0.9, -0.50006, 0.9, 0.01, 0.01, 0.01, (1, 0, 0, 0), (1.0, 0.0, 0.0))
    b.add(-0.6, -0.6, -0.500056, 0.9, 0.6, 0.6, 0.6, (1, 0, 0, 0), (0.0, 1.0, 0.0))
    spl = [(b.decode(0), np.eye(4), raster.SplatEdit())]
    plain = H.run_oracle(cam, s, spl)
    under_colour = np.empty((s.height, s.width, 4), np.float32)
    under_colour[...] = (0.2, 0.3, 0.4, 1.0)
    under_depth = np.full((s.height, s.width), plain["proj"][0, 2], np.float32)
    gpu = H.run_gpu_composite(cam, s, spl, [], under=(under_colour, under_depth), rast=rast)
    ora = H.run_oracle_composite(cam, s, spl, [], under_colour, under_depth)
    np.testing.assert_array_equal(gpu["colour"][16:32, 16:32], under_colour[16:32, 16:32])
    shown = np.any(gpu["colour"] != under_colour, axis=2)
    np.testing.assert_array_equal(shown, np.any(ora["colour"] != under_colour, axis=2))
    np.testing.assert_allclose(gpu["colour"], ora["colour"], rtol=0, atol=2e-5)


def test_argument_checks(rast):
    import torch
    print("composite argument checks", flush=True)
    s = raster.RenderSettings(width=32, height=32)
    cam = raster.Camera.looking_at((0.0, 0.0, 5.0), (0.0, 0.0, 0.0))
    targets = raster.RenderTargets(32, 32)
    pos, _ = F.random_points(10, 1)
    gp = raster.GpuPoints(torch.from_numpy(pos).cuda())
    with pytest.raises(LrtError) as err:   # no colours and no constant colour
        rast.render(cam, [], s, targets, points=[raster.PointInstance(gp)])
    assert err.value.code == 3
    with pytest.raises(LrtError) as err:   # an under-layer of another size
        rast.render(cam, [], s, targets, under=raster.RenderTargets(16, 32))
    assert err.value.code == 3
    with pytest.raises(LrtError) as err:
        rast.render(cam, [], s, targets, points=[raster.PointInstance(gp, style=raster.PointStyle(size_mode=2))])
    assert err.value.code == 3
    # a constant colour needs no colours; an empty frame draws the background
    rast.render(cam, [], s, targets, points=[raster.PointInstance(
        gp, style=raster.PointStyle(constant_colour=True, size_mode=1, size=4.0))])
    torch.cuda.synchronize()
    assert (targets.colour[..., 3] == 1.0).any()
    rast.render(cam, [], s, targets, points=[raster.PointInstance(None)])
    torch.cuda.synchronize()
    assert not targets.colour.any()
