"""The HIP levels of detail against the oracle on the same splats: the Morton
order, each level's groups and cells and every merged Gaussian bit for bit;
the cut's gathered cloud bit for bit over thresholds, eyes and a placed
instance; and the reference's own LOD tests (tests/lod/test_lod.cpp:82-185):
threshold 0 draws the cloud as it is (p99 <= 1), a one-splat cell merges into
its splat (p99 <= 1), and a dense cloud far away draws a fraction of its
splats within the stated image bounds."""
import numpy as np
import pytest

import fixtures as F
import harness as H
import oracle_lib as O
from lucabrtrender_amd import raster

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rast():
    return raster.TileRasterizer()


def _build(rast, scene, **settings):
    dev = scene.to_device()
    lod = rast.lod_build(dev, raster.LodBuildSettings(**settings))
    ora = O.lod_build(scene.positions, scene.shape, scene.sh, scene.rest_per_colour,
                      max_group_fraction=settings.get("max_group_fraction", 0.5),
                      coarsest_level=settings.get("coarsest_level", 1))
    return dev, lod, ora


def test_the_build_matches_the_oracle_bit_for_bit(rast):
    print("lod build 5000 splats", flush=True)
    scene = F.random_cloud(5000, 41).decode(3)
    _, lod, ora = _build(rast, scene)
    assert lod.count == 5000 and lod.merged == sum(len(lv["cells"]) for lv in ora["levels"])
    assert [lv["level"] for lv in lod.levels] == [lv["level"] for lv in ora["levels"]]
    np.testing.assert_array_equal(np.float32(lod.extent), ora["extent"])
    splats, groups = lod.splats()
    got = splats.read(rast)
    np.testing.assert_array_equal(got["positions"].view(np.uint32), ora["positions"].view(np.uint32))
    np.testing.assert_array_equal(got["shape"], ora["shape"])
    np.testing.assert_array_equal(got["sh"], ora["sh"])
    np.testing.assert_array_equal(groups, ora["groups"])
    for k, lv in enumerate(ora["levels"]):
        view, cells = lod.gaussians(k)
        g = view.read(rast)
        np.testing.assert_array_equal(cells, lv["cells"])
        np.testing.assert_array_equal(g["positions"].view(np.uint32), lv["positions"].view(np.uint32))
        np.testing.assert_array_equal(g["shape"], lv["shape"])
        np.testing.assert_array_equal(g["sh"], lv["sh"])


def _cut_both(rast, lod, ora, cam, settings, threshold, o2w=np.eye(4)):
    drawn, stats = rast.lod_select(cam, [raster.LodInstance(lod, tuple(o2w.reshape(-1)))], threshold, settings)
    proj = raster.projection_for(cam, settings.width, settings.height)
    inv = np.linalg.inv(o2w)
    eye = inv @ np.array([proj.eye_world[0], proj.eye_world[1], proj.eye_world[2], 1.0])
    scale = max(np.linalg.norm(o2w[:3, c]) for c in range(3))
    ppu = np.float32(proj.focal_x * scale if proj.orthographic else proj.focal_x)
    pos, shape, sh, merged, splats = O.lod_cut(ora, eye[:3].astype(np.float32), ppu, threshold,
                                               bool(proj.orthographic))
    return drawn, stats[0], (pos, shape, sh, merged, splats)


@pytest.mark.parametrize("view", ["near", "far", "inside", "ortho", "placed"])
def test_the_cut_matches_the_oracle_bit_for_bit(rast, view):
    print(f"lod cut {view}", flush=True)
    scene = F.random_cloud(5000, 41).decode(3)
    _, lod, ora = _build(rast, scene)
    s = raster.RenderSettings(width=200, height=150)
    o2w = np.eye(4)
    cam = raster.Camera.looking_at((1.5, 1.0, 6.0), (0.0, 0.0, 0.0))
    if view == "far":
        cam = raster.Camera.looking_at((6.0, 4.0, 30.0), (0.0, 0.0, 0.0))
    elif view == "inside":
        cam = raster.Camera.looking_at((0.1, 0.2, 0.3), (1.0, 0.0, 0.0))
    elif view == "ortho":
        cam.lens = raster.Lens(orthographic=True, focal=5.0)
    elif view == "placed":
        o2w[:3, :3] *= 1.7
        o2w[:3, 3] = (0.5, -0.3, -2.0)
    for threshold in (0.0, 1.0, 4.0, 16.0, 1e6):
        drawn, stats, (pos, shape, sh, merged, splats) = _cut_both(rast, lod, ora, cam, s, threshold, o2w)
        assert (stats.merged, stats.splats, stats.available) == (merged, splats, 5000), threshold
        assert len(drawn) == (1 if merged + splats else 0)
        if drawn:
            got = drawn[0].splats.read(rast)
            np.testing.assert_array_equal(got["positions"].view(np.uint32), pos.view(np.uint32))
            np.testing.assert_array_equal(got["shape"], shape)
            np.testing.assert_array_equal(got["sh"], sh)


def _render(rast, cam, instances, s):
    import torch
    t = raster.RenderTargets(s.width, s.height)
    rast.render(cam, instances, s, t)
    torch.cuda.synchronize()
    return t.colour.cpu().numpy()


def test_threshold_zero_draws_the_cloud_as_it_is(rast):
    """tests/lod/test_lod.cpp:82-105."""
    print("lod threshold 0", flush=True)
    scene = F.random_cloud(5000, 41).decode(3)
    dev, lod, _ = _build(rast, scene)
    s = raster.RenderSettings(width=200, height=150)
    cam = raster.Camera.looking_at((1.5, 1.0, 6.0), (0.0, 0.0, 0.0))
    full = _render(rast, cam, [raster.SplatInstance(dev)], s)
    drawn, stats = rast.lod_select(cam, [raster.LodInstance(lod)], 0.0, s)
    assert (stats[0].splats, stats[0].merged) == (5000, 0)
    cut = _render(rast, cam, drawn, s)
    p99, mx, over2 = O.compare_images(full, cut)
    assert p99 <= 1 and over2 * 100 <= s.width * s.height, (p99, mx, over2)


def test_a_cell_holding_one_splat_merges_into_that_splat(rast):
    """tests/lod/test_lod.cpp:107-153."""
    print("lod single-splat cells", flush=True)
    b = F.CloudBuilder()
    rng = F.Lcg(0)
    for p in [(-2, -2, -2), (2, -2, -2), (-2, 2, -2), (2, 2, 2), (-2, 2, 2), (2, -2, 2)]:
        opacity = rng.range(0.3, 0.95)
        sx, sy, sz = rng.range(0.1, 0.5), rng.range(0.1, 0.5), rng.range(0.1, 0.5)
        turn = [rng.range(-1, 1) for _ in range(4)]
        dc = [rng.range(0.1, 1.0) for _ in range(3)]
        b.add(p[0], p[1], p[2], opacity, sx, sy, sz, turn, dc, (0.2, -0.1, 0.3, 0.0, 0.1, -0.2, 0.1, 0.0, -0.3))
    scene = b.decode(3)
    dev, lod, _ = _build(rast, scene, max_group_fraction=1.0, coarsest_level=1)
    assert lod.levels[0]["groups"] == 6
    s = raster.RenderSettings(width=200, height=150)
    cam = raster.Camera.looking_at((3.0, 2.0, 12.0), (0.0, 0.0, 0.0))
    full = _render(rast, cam, [raster.SplatInstance(dev)], s)
    drawn, stats = rast.lod_select(cam, [raster.LodInstance(lod)], 1e6, s)
    assert (stats[0].splats, stats[0].merged) == (0, 6)
    cut = _render(rast, cam, drawn, s)
    p99, mx, over2 = O.compare_images(full, cut)
    assert p99 <= 1 and over2 * 100 <= s.width * s.height, (p99, mx, over2)


def test_far_away_a_dense_cloud_draws_a_fraction_and_looks_the_same(rast):
    """tests/lod/test_lod.cpp:155-185: bounds p99 <= 1 / 12 / 36 at 1 / 2 / 4 px."""
    print("lod far dense cloud", flush=True)
    scene = F.random_cloud(60000, 43, 0.003, 0.05).decode(3)
    dev, lod, _ = _build(rast, scene)
    s = raster.RenderSettings(width=240, height=180)
    cam = raster.Camera.looking_at((6.0, 4.0, 30.0), (0.0, 0.0, 0.0))
    full = _render(rast, cam, [raster.SplatInstance(dev)], s)
    for threshold, bound, share in ((1.0, 1, None), (2.0, 12, 0.75), (4.0, 36, 0.25)):
        drawn, stats = rast.lod_select(cam, [raster.LodInstance(lod)], threshold, s)
        cut = _render(rast, cam, drawn, s)
        p99, mx, over2 = O.compare_images(full, cut)
        shown = stats[0].splats + stats[0].merged
        print(f"  threshold {threshold}: {stats[0].splats} splats + {stats[0].merged} merged of "
              f"{stats[0].available}; p99 {p99}, max {mx}", flush=True)
        assert p99 <= bound
        if share is not None:
            assert shown < share * stats[0].available


def test_instances_with_their_own_thresholds(rast):
    print("lod two instances", flush=True)
    scene = F.random_cloud(5000, 41).decode(3)
    dev, lod, _ = _build(rast, scene)
    s = raster.RenderSettings(width=200, height=150)
    cam = raster.Camera.looking_at((1.5, 1.0, 9.0), (0.0, 0.0, 0.0))
    moved = np.eye(4)
    moved[:3, 3] = (3.0, 0.0, -3.0)
    instances = [raster.LodInstance(lod), raster.LodInstance(None),
                 raster.LodInstance(lod, tuple(moved.reshape(-1)), threshold=0.0)]
    drawn, stats = rast.lod_select(cam, instances, 1e6, s)
    assert len(drawn) == 2 and len(stats) == 3
    assert stats[0].splats == 0 and stats[0].merged == lod.levels[0]["groups"]
    assert (stats[1].splats, stats[1].merged, stats[1].available) == (0, 0, 0)
    assert (stats[2].splats, stats[2].merged) == (5000, 0)
    assert drawn[1].object_to_world == tuple(moved.reshape(-1))
    both = _render(rast, cam, drawn, s)
    alone = _render(rast, cam, [raster.SplatInstance(dev, tuple(moved.reshape(-1)))], s)
    assert O.compare_images(both, alone)[1] > 0   # the merged instance adds to the frame
