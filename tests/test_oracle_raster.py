"""Pins the oracle's tile rasteriser against its independent brute force, on
the reference's own scenes and tolerances (tests/render/test_splat_render.cpp):
randomCloud(3000, 7) at 250x190 over five cameras (p99 <= 2, max <= 16 8-bit
code values), the nearer splat winning from either side, antialiasing's energy,
and the binning invariants no retyped formula can satisfy by accident."""
import numpy as np
import pytest

import fixtures as F
import harness as H
import oracle_lib as O
from lucabrtrender_amd import raster

W, HGT = 250, 190


@pytest.fixture(scope="module")
def cloud():
    return F.random_cloud(3000, 7).decode(3)


def _cam(eye, target, **lens):
    c = raster.Camera.looking_at(eye, target)
    c.lens = raster.Lens(**lens)
    return c


SECTIONS = {
    "perspective": (_cam((1.5, 1.0, 6.0), (0, 0, 0), focal=30.0), dict(), True),
    "frustum-edge": (_cam((0.3, -0.2, 2.2), (0, 0, 0), focal=18.0), dict(), False),
    "orthographic": (_cam((0.0, 3.0, 5.0), (0, 0, 0), orthographic=True, focal=5.0), dict(), False),
    "no-aa-window": (_cam((-2.0, 0.5, 5.0), (0, 0, 0), window_translate=(0.2, 0.0), window_scale=(1.0, 1.3),
                          window_roll=25.0), dict(antialias=False), False),
}


def _two_instances(scene):
    o2w = O.lib  # noqa: F841  (documentation: translation(0.5,0,-1.5) * rotationY(40) * scaling(1,0.7,1.3))
    t = (O.dbl(np.eye(4)))
    import ctypes as C
    a, b, c_, ab, abc = ((C.c_double * 16)() for _ in range(5))
    O.lib.or_translation(O.dbl([0.5, 0.0, -1.5]), a)
    O.lib.or_rotation_y(40.0, b)
    O.lib.or_scaling(O.dbl([1.0, 0.7, 1.3]), c_)
    O.lib.or_mat_mul(a, b, ab)
    O.lib.or_mat_mul(ab, c_, abc)
    del t
    return [(scene, np.eye(4), raster.SplatEdit()), (scene, np.array(abc[:]).reshape(4, 4), raster.SplatEdit())]


@pytest.mark.parametrize("name", list(SECTIONS) + ["two-instances"])
def test_tile_raster_renders_what_the_brute_force_renders(cloud, name):
    if name == "two-instances":
        cam, extra, strict = _cam((2.0, 1.5, 7.0), (0.0, 0.0, -0.5)), dict(), False
        inst = _two_instances(cloud)
    else:
        cam, extra, strict = SECTIONS[name]
        inst = [(cloud, np.eye(4), raster.SplatEdit())]
    s = raster.RenderSettings(width=W, height=HGT, **extra)
    ours = H.run_oracle(cam, s, inst, threads=0)
    truth, _ = O.reference(ours["frames"], ours["clouds"], ours["n"], threads=0)
    p99, mx, over2 = O.compare_images(ours["colour"], truth)
    lin99, linmax = O.linear_error(ours["colour"], truth)
    print(f"{name}: {ours['pairs']} pairs: p99 {p99}, max {mx}, {over2} over 2; linear p99 {lin99:.2e} "
          f"max {linmax:.2e}")
    assert p99 <= 2
    if strict:
        assert mx <= 16
    # At the rasteriser's own tie rule (24-bit depth key, then index) the brute
    # force differs from the tile walk only by float rounding.
    keyed, _ = O.reference(ours["frames"], ours["clouds"], ours["n"], threads=0, keyed=True)
    k99, kmax = O.linear_error(ours["colour"], keyed)
    assert k99 <= 1e-5 and kmax <= 1e-3, (k99, kmax)


def test_fp32_brute_force_agrees_with_fp64(cloud):
    cam, _, _ = SECTIONS["perspective"]
    s = raster.RenderSettings(width=W, height=HGT)
    frames, clouds, n = H.oracle_frames(cam, s, [(cloud, np.eye(4), raster.SplatEdit())])
    c32, d32 = O.reference(frames, clouds, n)
    c64, d64 = O.reference(frames, clouds, n, f64=True)
    p99, mx = O.linear_error(c32, c64)
    print(f"fp32 vs fp64 brute force: p99 {p99:.2e}, max {mx:.2e}")
    assert p99 <= 1e-4 and mx <= 2e-2


def test_the_nearer_splat_wins_from_either_side_and_left_stays_left():
    b = F.CloudBuilder()
    b.add(0.0, 0.0, -1.0, 0.99, 0.5, 0.5, 0.05, (1, 0, 0, 0), (1, 0, 0))   # red, z = -1
    b.add(0.0, 0.0, 1.0, 0.99, 0.5, 0.5, 0.05, (1, 0, 0, 0), (0, 0, 1))    # blue, z = +1
    b.add(-3.0, 0.0, 0.0, 0.99, 0.3, 0.3, 0.3, (1, 0, 0, 0), (0, 1, 0))    # green, -x
    scene = b.decode(3)
    s = raster.RenderSettings(width=200, height=100, linearise=False)

    def render(eye):
        cam = raster.Camera.looking_at(eye, (0.0, 0.0, 0.0))
        return H.run_oracle(cam, s, [(scene, np.eye(4), raster.SplatEdit())])["colour"]

    front = render((0.0, 0.0, 10.0))
    assert front[50, 100, 2] > 0.9 and front[50, 100, 0] < 0.1
    assert int(np.argmax(front[50, :, 1])) < 100
    back = render((0.0, 0.0, -10.0))
    assert back[50, 100, 0] > 0.9 and back[50, 100, 2] < 0.1
    assert int(np.argmax(back[50, :, 1])) > 100


def test_antialiasing_pays_back_the_dilation_energy():
    b = F.CloudBuilder()
    b.add(0.0, 0.0, -50.0, 0.9, 0.002, 0.002, 0.002, (1, 0, 0, 0), (1, 1, 1))
    scene = b.decode(3)

    def energy(aa):
        s = raster.RenderSettings(width=64, height=64, antialias=aa)
        return float(H.run_oracle(raster.Camera(), s, [(scene, np.eye(4), raster.SplatEdit())])["colour"][..., 3].sum())

    without = energy(False)
    assert without > 0.0 and energy(True) < 0.6 * without


@pytest.mark.parametrize("tile", [16, 8, 32])
def test_binning_invariants(cloud, tile):
    """Every pixel centre a splat reaches (its power at that centre, in double,
    below its cutoff) lies in a tile whose range holds the splat; each tile's run
    holds only its own pairs, in (24-bit depth key, slot) order. At the
    reference's 16-pixel tiles and at this library's 8 and 32."""
    cam, _, _ = SECTIONS["perspective"]
    s = raster.RenderSettings(width=W, height=HGT, tile_size=tile)
    r = H.run_oracle(cam, s, [(cloud, np.eye(4), raster.SplatEdit())])
    f = r["frames"][0]
    tiles, splats, ranges, proj = r["pair_tiles"], r["pair_splats"], r["ranges"], r["proj"]
    assert np.all(np.diff(tiles.astype(np.int64)) >= 0)
    for t in range(f.tiles_x * f.tiles_y):
        a, e = ranges[t]
        assert np.all(tiles[a:e] == t)
        keys = (proj[splats[a:e], 2].view(np.uint32) >> 7).astype(np.int64)
        order = keys * (1 << 32) + splats[a:e].astype(np.int64)
        assert np.all(np.diff(order) > 0)
    assert int((ranges[:, 1] - ranges[:, 0]).sum()) == r["pairs"]
    members = {}
    for k in range(r["pairs"]):
        members.setdefault(int(tiles[k]), set()).add(int(splats[k]))
    ys, xs = np.mgrid[0:HGT, 0:W]
    px, py = xs.reshape(-1) + 0.5, ys.reshape(-1) + 0.5
    assert f.tiles_x == -(-W // tile) and f.tiles_y == -(-HGT // tile)
    tile_of = (ys.reshape(-1) // tile) * f.tiles_x + xs.reshape(-1) // tile
    for i in np.nonzero(r["touched"])[0][:600]:
        sc, co = proj[i, :4].astype(np.float64), proj[i, 4:8].astype(np.float64)
        dx, dy = px - sc[0], py - sc[1]
        power = co[0] * dx * dx + 2 * co[1] * dx * dy + co[2] * dy * dy
        reached = np.unique(tile_of[power <= sc[3] * (1 - 1e-6)])
        for t in reached:
            assert i in members.get(int(t), ()), (i, t)


@pytest.mark.parametrize("name", list(SECTIONS))
def test_tile_size_changes_the_pairs_not_the_image(cloud, name):
    """Every splat reaching a pixel is in its tile's list at any tile size and
    each list walks in (depth key, index) order, so 8- and 32-pixel tiles draw
    the 16-pixel frame bit for bit, from about 3x and 0.4x its pairs."""
    cam, extra, _ = SECTIONS[name]
    inst = [(cloud, np.eye(4), raster.SplatEdit())]
    runs = {t: H.run_oracle(cam, raster.RenderSettings(width=W, height=HGT, tile_size=t, **extra), inst)
            for t in (16, 8, 32)}
    for t in (8, 32):
        np.testing.assert_array_equal(runs[t]["colour"].view(np.uint32), runs[16]["colour"].view(np.uint32))
        np.testing.assert_array_equal(runs[t]["depth"].view(np.uint32), runs[16]["depth"].view(np.uint32))
    assert runs[8]["pairs"] > 2 * runs[16]["pairs"] > 4 * runs[32]["pairs"]


def test_splat_edit_renders_as_the_brute_force_renders_it(cloud):
    cam, _, _ = SECTIONS["perspective"]
    s = raster.RenderSettings(width=W, height=HGT)
    for ed in [raster.SplatEdit(active=True, shape=0, mode=0, size=(1.0, 0.8, 1.2)),
               raster.SplatEdit(active=True, shape=1, mode=1, size=(0.9, 0, 0), centre=(0.3, 0.0, 0.2)),
               raster.SplatEdit(active=True, mode=2, tint=(1.2, 0.8, 0.6), saturation=0.3, brightness=1.4,
                                opacity=0.7, size=(1.5, 1.5, 1.5)),
               raster.SplatEdit(active=True, mode=2, min_opacity=0.5, max_scale=0.3, invert=True)]:
        ours = H.run_oracle(cam, s, [(cloud, np.eye(4), ed)])
        truth, _ = O.reference(ours["frames"], ours["clouds"], ours["n"])
        p99, mx, _ = O.compare_images(ours["colour"], truth)
        assert p99 <= 2, (ed, p99, mx)
