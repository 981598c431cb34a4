"""Pins the oracle's disc route and composite blend (pointsProject,
shaders/lrt/splat/points_project.slang:33-86; splatBlendComposite,
splat_blend.slang:23-120):

- one disc against its closed form: the pixels whose centres lie within the
  radius (size / 2 * focal / z, size / 2 in pixels, never below half a pixel)
  take its colour and depth, the rest the background;
- discs against the oracle's brute force, which projects them on its own
  (centre by division, no tiles) and sorts by exact depth (the reference's
  tests/render/test_points.cpp:75-99, p99 <= 2 code values);
- splats over a disc layer given as the under-layer against splats and discs in
  one sort (test_points.cpp:136-197 with the disc route drawing the layer);
- the under-layer's arithmetic: what lies behind is composited with the
  transmittance left, bit for bit against the splat-only frame;
- the cut: the first pair of a tile at or behind a pixel's under depth ends
  its walk even where that pair does not reach the pixel (splat_blend.slang:52-54);
- each disc's tile count against a float64 count of the tiles whose pixel
  centres it covers.
"""
import numpy as np
import pytest

import fixtures as F
import harness as H
import oracle_lib as O
from lucabrtrender_amd import raster

W, HGT = 230, 170


def _points_cam():
    return raster.Camera.looking_at((1.0, 2.0, 6.0), (0.0, 0.0, 0.0))


@pytest.mark.parametrize("mode,size", [(0, 0.4), (1, 9.0), (0, 1e-5)])
def test_one_disc_against_its_closed_form(mode, size):
    cam = raster.Camera.looking_at((0.0, 0.0, 5.0), (0.0, 0.0, 0.0))
    s = raster.RenderSettings(width=65, height=49, linearise=False, background=(0.1, 0.2, 0.3, 0.5))
    positions = np.array([[0.0, 0.0, 0.0, 1.0]], np.float32)
    colours = F.pack_point_colours(np.array([[0.25, 0.5, 0.75]], np.float32))
    style = raster.PointStyle(size_mode=mode, size=size)
    out = H.run_oracle_composite(cam, s, [], [(positions, colours, np.eye(4), style)])
    proj = O.projection_for(H.oracle_camera(cam), s.width, s.height)
    z = 5.0
    radius = size * 0.5 if mode == 1 else size * 0.5 * proj.focal_x / z
    radius = max(radius, 0.5)
    ys, xs = np.mgrid[0:s.height, 0:s.width]
    d2 = (xs + 0.5 - proj.centre_x) ** 2 + (ys + 0.5 - proj.centre_y) ** 2
    inside = d2 <= radius * radius
    clear = np.abs(d2 / (radius * radius) - 1.0) > 1e-4   # leave the float edge alone
    assert inside.any()
    np.testing.assert_array_equal(out["colour"][inside & clear], [[0.25, 0.5, 0.75, 1.0]] * int((inside & clear).sum()))
    np.testing.assert_allclose(out["depth"][inside & clear], z, rtol=1e-6)
    outside = ~inside & clear
    np.testing.assert_array_equal(out["colour"][outside],
                                  np.broadcast_to(np.float32([0.1, 0.2, 0.3, 0.5]), (int(outside.sum()), 4)))
    np.testing.assert_array_equal(out["depth"][outside], 0.0)
    assert out["visible"] == 1
    rec = out["proj"][0]
    # cutoff 1, alpha 1, colour.w 1 (opaque); the conic 1 / r^2 where r is exact in float
    assert rec[3] == 1.0 and rec[7] == 1.0 and rec[11] == 1.0 and rec[5] == 0.0
    if mode == 1 or radius == 0.5:
        r32 = np.float32(radius)
        assert rec[4] == rec[6] == np.float32(1.0) / (r32 * r32)


@pytest.mark.parametrize("mode,size", [(0, 0.06), (1, 5.0)])
def test_discs_match_the_brute_force(mode, size):
    cam = _points_cam()
    s = raster.RenderSettings(width=W, height=HGT, linearise=False)
    pos, col = F.random_points(4000, 3)
    pts = [(pos, col, np.eye(4), raster.PointStyle(size_mode=mode, size=size))]
    ours = H.run_oracle_composite(cam, s, [], pts)
    truth, _ = H.run_reference_points(cam, s, [], pts)
    p99, mx, over2 = O.compare_images(ours["colour"], truth)
    assert p99 <= 2, (p99, mx, over2)
    assert ours["visible"] > 1000


def test_constant_colour_and_a_placed_instance():
    cam = _points_cam()
    s = raster.RenderSettings(width=120, height=90, linearise=False)
    pos, _ = F.random_points(600, 9)
    o2w = np.eye(4)
    o2w[:3, 3] = (0.3, -0.2, 0.5)
    style = raster.PointStyle(size_mode=1, size=4.0, constant_colour=True, colour=(0.9, 0.1, 0.4))
    pts = [(pos, None, o2w, style)]
    ours = H.run_oracle_composite(cam, s, [], pts)
    truth, _ = H.run_reference_points(cam, s, [], pts)
    p99, mx, _ = O.compare_images(ours["colour"], truth)
    assert p99 <= 2
    covered = ours["colour"][..., 3] == 1.0
    assert covered.sum() > 500
    np.testing.assert_array_equal(ours["colour"][covered][:, :3],
                                  np.broadcast_to(np.float32([0.9, 0.1, 0.4]), (int(covered.sum()), 3)))


def test_splats_over_a_disc_layer_match_splats_and_discs_in_one_sort():
    cam = raster.Camera.looking_at((0.5, 0.5, 7.0), (0.0, 0.0, 0.0))
    s = raster.RenderSettings(width=256, height=180)
    pos, col = F.random_points(8000, 5)
    fog = F.decode_encoded(F.fog_records(), F.FOG_ENCODING)
    pts = [(pos, col, np.eye(4), raster.PointStyle(size=0.05))]
    spl = [(fog, np.eye(4), raster.SplatEdit())]
    layer = H.run_oracle_composite(cam, s, [], pts)
    over = H.run_oracle_composite(cam, s, spl, [], layer["colour"], layer["depth"])
    mixed = H.run_oracle_composite(cam, s, spl, pts)
    p99, mx, over2 = O.compare_images(over["colour"], mixed["colour"])
    assert p99 <= 2, (p99, mx, over2)
    # and the splats do change the image: the fog is not behind everything
    only = H.run_oracle_composite(cam, s, [], pts)
    assert O.compare_images(only["colour"], mixed["colour"])[1] > 10


def test_the_under_layer_is_composited_with_the_transmittance_left():
    cam = raster.Camera.looking_at((1.5, 1.0, 6.0), (0.0, 0.0, 0.0))
    s = raster.RenderSettings(width=96, height=80)
    scene = F.random_cloud(400, 7).decode(3)
    spl = [(scene, np.eye(4), raster.SplatEdit())]
    under_colour = np.empty((s.height, s.width, 4), np.float32)
    under_colour[...] = (0.2, 0.3, 0.4, 1.0)
    far = np.full((s.height, s.width), 1000.0, np.float32)
    plain = H.run_oracle(cam, s, spl)
    over = H.run_oracle_composite(cam, s, spl, [], under_colour, far)
    t = plain["final_t"]
    want = plain["colour"] + t[..., None] * under_colour   # the splat-only frame has a zero background
    np.testing.assert_array_equal(over["colour"], want.astype(np.float32))
    cov = 1.0 - t
    want_depth = np.where(cov > 1e-6, plain["depth"] * cov, 0.0) + t * 1000.0
    np.testing.assert_allclose(over["depth"], want_depth, rtol=2e-5, atol=1e-4)

    # an under depth nearer than every splat shows the layer alone
    near = np.full((s.height, s.width), 0.5, np.float32)
    hidden = H.run_oracle_composite(cam, s, spl, [], under_colour, near)
    np.testing.assert_array_equal(hidden["colour"], under_colour)
    np.testing.assert_array_equal(hidden["depth"], near)

    # an under depth of 0 is no under-layer: the background lies behind
    none = np.zeros((s.height, s.width), np.float32)
    bare = H.run_oracle_composite(cam, s, spl, [], under_colour, none)
    np.testing.assert_array_equal(bare["colour"], plain["colour"])


def _cut_scene():
    """Two splats in one tile with the same 24-bit depth key: a tiny one (slot 0)
    a few ulps behind a wide one (slot 1), so the wide one comes second."""
    cam = raster.Camera.looking_at((0.0, 0.0, 10.0), (0.0, 0.0, 0.0))
    cam.lens = raster.Lens(orthographic=True, focal=5.0)
    s = raster.RenderSettings(width=64, height=64, linearise=False)
    b = F.CloudBuilder()
    b.add(-0.9, -0.9, -0.50006, 0.9, 0.01, 0.01, 0.01, (1, 0, 0, 0), (1.0, 0.0, 0.0))
    b.add(-0.6, -0.6, -0.500056, 0.9, 0.6, 0.6, 0.6, (1, 0, 0, 0), (0.0, 1.0, 0.0))
    return cam, s, [(b.decode(0), np.eye(4), raster.SplatEdit())]


def test_the_first_pair_at_or_behind_the_under_depth_ends_the_walk():
    cam, s, spl = _cut_scene()
    plain = H.run_oracle(cam, s, spl)
    p = plain["proj"]
    assert plain["depth_keys"][0] == plain["depth_keys"][1] and p[1, 2] < p[0, 2]
    tile = lambda x, y: (int(y) // 16) * 4 + int(x) // 16   # noqa: E731
    assert tile(p[0, 0], p[0, 1]) == tile(p[1, 0], p[1, 1]) == 5
    under_colour = np.empty((s.height, s.width, 4), np.float32)
    under_colour[...] = (0.2, 0.3, 0.4, 1.0)
    under_depth = np.full((s.height, s.width), p[0, 2], np.float32)   # the tiny splat's depth exactly
    out = H.run_oracle_composite(cam, s, spl, [], under_colour, under_depth)
    shown = np.any(out["colour"] != under_colour, axis=2)
    # tile (1, 1) (pixels 16-31): the tiny splat comes first and ends every walk there
    assert not shown[16:32, 16:32].any()
    # elsewhere the wide splat, nearer than the layer, is drawn
    assert shown.sum() > 500
    assert shown[0:16].any() and shown[32:48].any()


@pytest.mark.parametrize("mode,size", [(0, 0.3), (1, 37.0), (1, 0.2)])
def test_disc_tile_counts_are_the_tiles_whose_pixel_centre_box_they_reach(mode, size):
    """frame.slang:133-138: a tile counts when the disc reaches some point of the
    box spanned by its first and last pixel centres; counted here in float64
    from the nearest point of each box."""
    cam = _points_cam()
    s = raster.RenderSettings(width=100, height=70, linearise=False)
    pos, col = F.random_points(300, 17)
    out = H.run_oracle_composite(cam, s, [], [(pos, col, np.eye(4), raster.PointStyle(size_mode=mode, size=size))])
    proj = out["proj"]
    tx, ty = np.meshgrid(np.arange(7), np.arange(5))
    lox, loy = tx * 16 + 0.5, ty * 16 + 0.5
    hix = np.minimum((tx + 1) * 16, s.width) - 0.5
    hiy = np.minimum((ty + 1) * 16, s.height) - 0.5
    checked = 0
    for i in range(proj.shape[0]):
        rec = proj[i].astype(np.float64)
        if out["touched"][i] == 0:
            continue
        dx = np.clip(rec[0], lox, hix) - rec[0]
        dy = np.clip(rec[1], loy, hiy) - rec[1]
        power = rec[4] * dx * dx + rec[6] * dy * dy
        if (np.abs(power - 1.0) < 1e-4).any():
            continue   # a box on the float edge: either count is right
        assert out["touched"][i] == int((power <= 1.0).sum()), i
        checked += 1
    assert checked > 50
