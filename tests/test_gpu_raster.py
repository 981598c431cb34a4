"""The HIP rasteriser against the oracle on the same packed splats: counts,
projected records, sorted (tile, splat) pairs and tile ranges bit-exact; images
within tolerance of the oracle's restatement of the reference blend and of the
brute force, on the reference's own scenes and cameras
(tests/render/test_splat_render.cpp:70-121) and on the benchmark scenes."""
import numpy as np
import pytest

import fixtures as F
import harness as H
import oracle_lib as O
from lucabrtrender_amd import _abi, raster, scenes

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


CAMERAS = {
    "perspective": (_cam((1.5, 1.0, 6.0), (0, 0, 0), focal=30.0), dict()),
    "frustum-edge": (_cam((0.3, -0.2, 2.2), (0, 0, 0), focal=18.0), dict()),
    "orthographic": (_cam((0.0, 3.0, 5.0), (0, 0, 0), orthographic=True, focal=5.0), dict()),
    "no-aa-window": (_cam((-2.0, 0.5, 5.0), (0, 0, 0), window_translate=(0.2, 0.0), window_scale=(1.0, 1.3),
                          window_roll=25.0), dict(antialias=False)),
}


def _image_checks(gpu, ora, truth=None, tol_max=1e-2):
    d = np.abs(gpu["colour"].astype(np.float64) - ora["colour"])
    assert d.max() <= tol_max and np.quantile(d, 0.99) <= 1e-4, (d.max(), np.quantile(d, 0.99))
    dd = np.abs(gpu["depth"].astype(np.float64) - ora["depth"])
    assert np.quantile(dd, 0.99) <= 1e-4 * max(1.0, float(np.abs(ora["depth"]).max()))
    if truth is not None:
        p99, mx, _ = O.compare_images(gpu["colour"], truth)
        assert p99 <= 2, (p99, mx)


@pytest.mark.parametrize("name", list(CAMERAS))
def test_reference_scene_and_cameras(rast, cloud, name):
    cam, extra = CAMERAS[name]
    s = raster.RenderSettings(width=W, height=HGT, **extra)
    inst = [(cloud, np.eye(4), raster.SplatEdit())]
    gpu = H.run_gpu(cam, s, inst, rast)
    ora = H.run_oracle(cam, s, inst)
    H.assert_bit_exact(gpu, ora)
    truth, _ = O.reference(ora["frames"], ora["clouds"], ora["n"])
    _image_checks(gpu, ora, truth)
    # The walk's end agrees wherever no alpha or transmittance sits on a threshold
    # within the hardware exp2's last bits.
    assert np.mean(gpu["last_count"] == ora["last_count"]) >= 0.999


def test_two_instances_one_turned_scaled_and_moved(rast, cloud):
    o2w = np.eye(4)
    a = np.radians(40.0)
    o2w[:3, :3] = np.array([[np.cos(a), 0, np.sin(a)], [0, 1, 0], [-np.sin(a), 0, np.cos(a)]]) @ np.diag([1, 0.7, 1.3])
    o2w[:3, 3] = [0.5, 0.0, -1.5]
    inst = [(cloud, np.eye(4), raster.SplatEdit()), (cloud, o2w, raster.SplatEdit())]
    cam = _cam((2.0, 1.5, 7.0), (0.0, 0.0, -0.5))
    s = raster.RenderSettings(width=W, height=HGT)
    gpu = H.run_gpu(cam, s, inst, rast)
    ora = H.run_oracle(cam, s, inst)
    H.assert_bit_exact(gpu, ora)
    _image_checks(gpu, ora)


@pytest.mark.parametrize("edit", [
    raster.SplatEdit(active=True, shape=0, mode=0, size=(1.0, 0.8, 1.2)),
    raster.SplatEdit(active=True, shape=1, mode=1, size=(0.9, 0, 0), centre=(0.3, 0.0, 0.2)),
    raster.SplatEdit(active=True, mode=2, tint=(1.2, 0.8, 0.6), saturation=0.3, brightness=1.4, opacity=0.7,
                     size=(1.5, 1.5, 1.5)),
    raster.SplatEdit(active=True, mode=2, min_opacity=0.5, max_scale=0.3, invert=True),
])
def test_splat_edits(rast, cloud, edit):
    cam, _ = CAMERAS["perspective"]
    s = raster.RenderSettings(width=W, height=HGT)
    inst = [(cloud, np.eye(4), edit)]
    gpu = H.run_gpu(cam, s, inst, rast)
    ora = H.run_oracle(cam, s, inst)
    H.assert_bit_exact(gpu, ora)
    _image_checks(gpu, ora)


@pytest.mark.parametrize("settings", [
    dict(depth_mode=1, depth_threshold=0.3, background=(0.2, 0.3, 0.4, 1.0)),
    dict(linearise=False, max_sh_degree=1),
    dict(max_sh_degree=0, dilation=0.1),
])
def test_settings_variants(rast, cloud, settings):
    cam, _ = CAMERAS["frustum-edge"]
    s = raster.RenderSettings(width=W, height=HGT, **settings)
    inst = [(cloud, np.eye(4), raster.SplatEdit())]
    gpu = H.run_gpu(cam, s, inst, rast)
    ora = H.run_oracle(cam, s, inst)
    H.assert_bit_exact(gpu, ora)
    _image_checks(gpu, ora)


def test_plain_65k_full_size(rast):
    cfg = scenes.workload("plain-fwd-65k-640x480")
    s = raster.RenderSettings(width=cfg.width, height=cfg.height)
    inst = [(cfg.scene, np.eye(4), raster.SplatEdit())]
    gpu = H.run_gpu(cfg.camera, s, inst, rast)
    ora = H.run_oracle(cfg.camera, s, inst)
    print("plain 65k: pairs", ora["pairs"], "visible", ora["visible"])
    H.assert_bit_exact(gpu, ora)
    _image_checks(gpu, ora)
    # brute force on a pixel subset, exact view-z order and the rasteriser's key order
    rng = np.random.default_rng(0)
    xs, ys = rng.integers(0, cfg.width, 400), rng.integers(0, cfg.height, 400)
    exact, _ = O.reference(ora["frames"], ora["clouds"], ora["n"], pixels=(xs, ys))
    keyed, _ = O.reference(ora["frames"], ora["clouds"], ora["n"], pixels=(xs, ys), keyed=True)
    ours = gpu["colour"][ys, xs]
    e99, emax = O.linear_error(ours, exact)
    k99, kmax = O.linear_error(ours, keyed)
    print(f"plain 65k vs brute force: exact order p99 {e99:.2e} max {emax:.2e}; key order p99 {k99:.2e} "
          f"max {kmax:.2e}")
    assert e99 <= 1e-3 and k99 <= 1e-3 and kmax <= 1e-2


@pytest.mark.parametrize("count", [200_000])
def test_clustered_scene_scaled(rast, count):
    cfg = scenes.workload("scene-fwd-3M-1080p", count=count)
    s = raster.RenderSettings(width=cfg.width, height=cfg.height, dilation=cfg.dilation, filter3d=cfg.filter3d)
    inst = [(cfg.scene, np.eye(4), raster.SplatEdit())]
    gpu = H.run_gpu(cfg.camera, s, inst, rast)
    ora = H.run_oracle(cfg.camera, s, inst)
    H.assert_bit_exact(gpu, ora)
    _image_checks(gpu, ora)


def test_sort_stress_large_footprints_scaled(rast):
    cfg = scenes.workload("sort-stress-6M-2160p", count=60_000)
    s = raster.RenderSettings(width=cfg.width, height=cfg.height)
    inst = [(cfg.scene, np.eye(4), raster.SplatEdit())]
    gpu = H.run_gpu(cfg.camera, s, inst, rast)
    ora = H.run_oracle(cfg.camera, s, inst)
    t = ora["touched"]
    print("sort stress scaled: pairs", ora["pairs"], "max tiles a splat", int(t.max()))
    assert int(t.max()) > 256
    H.assert_bit_exact(gpu, ora)
    _image_checks(gpu, ora)


def test_clustered_3m_full_size_bit_exact_and_deterministic(rast):
    cfg = scenes.workload("scene-fwd-3M-1080p")
    s = raster.RenderSettings(width=cfg.width, height=cfg.height, dilation=cfg.dilation, filter3d=cfg.filter3d)
    inst = [(cfg.scene, np.eye(4), raster.SplatEdit())]
    gpu = H.run_gpu(cfg.camera, s, inst, rast)
    ora = H.run_oracle(cfg.camera, s, inst)
    print("scene 3M: pairs", ora["pairs"], "visible", ora["visible"])
    H.assert_bit_exact(gpu, ora)
    _image_checks(gpu, ora)
    again = H.run_gpu(cfg.camera, s, inst, rast)
    assert np.array_equal(again["colour"].view(np.uint32), gpu["colour"].view(np.uint32))
    assert np.array_equal(again["keys"], gpu["keys"])


def test_empty_and_degenerate_frames(rast, cloud):
    s = raster.RenderSettings(width=37, height=21, background=(0.1, 0.2, 0.3, 0.4))
    behind = _cam((0.0, 0.0, 50.0), (0.0, 0.0, 100.0))      # looking away: everything culled
    gpu = H.run_gpu(behind, s, [(cloud, np.eye(4), raster.SplatEdit())], rast)
    assert gpu["counts"][2] == 0
    assert np.allclose(gpu["colour"], np.array([0.1, 0.2, 0.3, 0.4], np.float32))
    one = F.CloudBuilder()
    one.add(0.0, 0.0, -3.0, 0.8, 0.1, 0.1, 0.1, (1, 0, 0, 0), (0.3, 0.6, 0.9))
    sc = one.decode(3)
    s1 = raster.RenderSettings(width=64, height=48)
    gpu = H.run_gpu(raster.Camera(), s1, [(sc, np.eye(4), raster.SplatEdit())], rast)
    ora = H.run_oracle(raster.Camera(), s1, [(sc, np.eye(4), raster.SplatEdit())])
    H.assert_bit_exact(gpu, ora)
    _image_checks(gpu, ora)


def test_errors_carry_the_reference_codes(rast, cloud):
    import torch
    dev = cloud.to_device()
    t = raster.RenderTargets(10, 10)
    with pytest.raises(raster.LrtError) as e:
        rast.render(raster.Camera(), [raster.SplatInstance(dev)], raster.RenderSettings(width=0, height=10), t)
    assert e.value.name == "InvalidArgument"
    with pytest.raises(raster.LrtError) as e:
        rast.render(raster.Camera(), [raster.SplatInstance(dev)], raster.RenderSettings(width=12, height=10), t)
    assert e.value.name == "InvalidArgument"
    with pytest.raises(raster.LrtError) as e:
        rast.render(raster.Camera(), [raster.SplatInstance(dev)],
                    raster.RenderSettings(width=10, height=10, filter3d=True), t)
    assert e.value.name == "InvalidArgument"
    bad = raster.GpuSplats(dev.positions, dev.shape, dev.sh, 3)
    bad._c.rest_per_colour = 5
    with pytest.raises(raster.LrtError) as e:
        rast.render(raster.Camera(), [raster.SplatInstance(bad)], raster.RenderSettings(width=10, height=10), t)
    assert e.value.name == "InvalidArgument"
    with pytest.raises(raster.LrtError) as e:
        rast.radix_sort(torch.zeros(4, dtype=torch.int64, device="cuda"), torch.zeros(4, dtype=torch.int32,
                        device="cuda"), torch.zeros(4, dtype=torch.int64, device="cuda"),
                        torch.zeros(4, dtype=torch.int32, device="cuda"), 65)
    assert e.value.name == "InvalidArgument"
    # still renders after the failures
    s = raster.RenderSettings(width=W, height=HGT)
    gpu = H.run_gpu(CAMERAS["perspective"][0], s, [(cloud, np.eye(4), raster.SplatEdit())], rast)
    assert gpu["counts"][2] > 0
    assert rast.debug(_abi.DEBUG_COUNTS)[2] == gpu["counts"][2]


@pytest.mark.parametrize("batch,order", [(64, 0), (64, 1), (128, 0), (128, 1), (256, 0), (256, 1)])
def test_every_batch_depth_and_tile_order_draws_the_same_frame(rast, cloud, batch, order):
    """Batch depth and launch order are performance parameters only."""
    cam, _ = CAMERAS["perspective"]
    s = raster.RenderSettings(width=W, height=HGT, blend_batch=batch, tile_order=order)
    inst = [(cloud, np.eye(4), raster.SplatEdit())]
    gpu = H.run_gpu(cam, s, inst, rast)
    ora = H.run_oracle(cam, s, inst)
    H.assert_bit_exact(gpu, ora)
    _image_checks(gpu, ora)
    cfg = scenes.workload("plain-fwd-65k-640x480", count=20000)
    s1 = raster.RenderSettings(width=cfg.width, height=cfg.height, blend_batch=batch, tile_order=order, depth_mode=1)
    inst1 = [(cfg.scene, np.eye(4), raster.SplatEdit())]
    gpu = H.run_gpu(cfg.camera, s1, inst1, rast)
    ora = H.run_oracle(cfg.camera, s1, inst1)
    H.assert_bit_exact(gpu, ora)
    _image_checks(gpu, ora)


@pytest.mark.parametrize("tile", [8, 32])
@pytest.mark.parametrize("name", ["perspective", "frustum-edge", "no-aa-window"])
def test_tile_sizes_bin_as_the_oracle_and_draw_the_16_pixel_frame(rast, cloud, name, tile):
    """At 8- and 32-pixel tiles the counts, records, pairs and ranges are the
    oracle's at that tile size, bit for bit, and the images are the 16-pixel
    frame's: the quadrants are the same 8x8 blocks and each walks the splats
    that reach it in the same (depth key, index) order."""
    print(f"tile size {tile} camera {name}", flush=True)
    cam, extra = CAMERAS[name]
    inst = [(cloud, np.eye(4), raster.SplatEdit())]
    base = H.run_gpu(cam, raster.RenderSettings(width=W, height=HGT, **extra), inst, rast)
    s = raster.RenderSettings(width=W, height=HGT, tile_size=tile, **extra)
    gpu = H.run_gpu(cam, s, inst, rast)
    ora = H.run_oracle(cam, s, inst)
    assert len(ora["ranges"]) == (-(-W // tile)) * (-(-HGT // tile))
    H.assert_bit_exact(gpu, ora)
    _image_checks(gpu, ora)
    np.testing.assert_array_equal(gpu["colour"].view(np.uint32), base["colour"].view(np.uint32))
    np.testing.assert_array_equal(gpu["depth"].view(np.uint32), base["depth"].view(np.uint32))
    np.testing.assert_array_equal(gpu["final_t"].view(np.uint32), base["final_t"].view(np.uint32))
    # The 1080p benchmark camera over a clustered scene, heavy tiles split.
    cfg = scenes.workload("plain-fwd-65k-640x480", count=20000)
    inst1 = [(cfg.scene, np.eye(4), raster.SplatEdit())]
    for segment in (0, 64):
        s1 = raster.RenderSettings(width=cfg.width, height=cfg.height, tile_size=tile, blend_segment=segment,
                                   depth_mode=1)
        gpu = H.run_gpu(cfg.camera, s1, inst1, rast)
        ora = H.run_oracle(cfg.camera, s1, inst1)
        H.assert_bit_exact(gpu, ora)
        _image_checks(gpu, ora)


@pytest.mark.parametrize("radix", [8, 9, 10])
@pytest.mark.parametrize("tile", [8, 16])
def test_sort_digit_widths_give_the_oracle_s_pairs(rast, radix, tile):
    """The frame's sort at digits of up to 8, 9 and 10 bits: the same pairs and
    ranges as the oracle's, bit for bit, in the passes the widths make."""
    print(f"radix {radix} tile {tile}", flush=True)
    cfg = scenes.workload("plain-fwd-65k-640x480", count=20000)
    inst = [(cfg.scene, np.eye(4), raster.SplatEdit())]
    s = raster.RenderSettings(width=cfg.width, height=cfg.height, tile_size=tile, radix_bits=radix)
    gpu = H.run_gpu(cfg.camera, s, inst, rast)
    ora = H.run_oracle(cfg.camera, s, inst)
    H.assert_bit_exact(gpu, ora)
    _image_checks(gpu, ora)


def _heaviest_first(ranges):
    """Tiles by the bit length of their pair count, descending, index order within one length."""
    counts = (ranges[:, 1].astype(np.int64) - ranges[:, 0].astype(np.int64))
    lengths = np.array([int(c).bit_length() for c in counts])
    return np.array(sorted(range(len(counts)), key=lambda t: (-lengths[t], t)), dtype=np.uint32)


@pytest.mark.gpu
@pytest.mark.parametrize("size", [(W, HGT), (640, 480), (1920, 1080)])
def test_tile_order_is_a_stable_heaviest_first_permutation(rast, cloud, size):
    """Every tile exactly once, buckets by pair-count bit length descending, stable inside a bucket."""
    cam, _ = CAMERAS["perspective"]
    s = raster.RenderSettings(width=size[0], height=size[1], tile_order=1)
    inst = [(cloud, np.eye(4), raster.SplatEdit())]
    H.run_gpu(cam, s, inst, rast)
    ranges = rast.debug(_abi.DEBUG_RANGES).reshape(-1, 2)
    order = rast.debug(_abi.DEBUG_TILE_ORDER)
    assert len(order) == len(ranges)
    np.testing.assert_array_equal(order, _heaviest_first(ranges))
    s.tile_order = 0
    H.run_gpu(cam, s, inst, rast)
    assert len(rast.debug(_abi.DEBUG_TILE_ORDER)) == 0


@pytest.mark.parametrize("segment", [16, 64, 300])
@pytest.mark.parametrize("depth_mode", [0, 1])
def test_split_tiles_draw_the_frame_the_one_walk_draws(rast, segment, depth_mode):
    """Tiles of more than `segment` pairs walked in segments by waves of their
    own and composited: the same images (to float association), the same last
    contributor and final transmittance, as the one walk; and the oracle's frame."""
    print(f"split blend segment {segment} depth mode {depth_mode}", flush=True)
    cfg = scenes.workload("plain-fwd-65k-640x480", count=20000)
    inst = [(cfg.scene, np.eye(4), raster.SplatEdit())]
    whole = H.run_gpu(cfg.camera, raster.RenderSettings(width=cfg.width, height=cfg.height, depth_mode=depth_mode,
                                                        blend_segment=0), inst, rast)
    s = raster.RenderSettings(width=cfg.width, height=cfg.height, depth_mode=depth_mode, blend_segment=segment)
    gpu = H.run_gpu(cfg.camera, s, inst, rast)
    split = int(np.sum(np.diff(gpu["ranges"], axis=1) > segment))
    assert split > 0
    np.testing.assert_allclose(gpu["colour"], whole["colour"], rtol=0, atol=3e-5)
    np.testing.assert_allclose(gpu["depth"], whole["depth"], rtol=1e-4, atol=1e-4)
    np.testing.assert_allclose(gpu["final_t"], whole["final_t"], rtol=1e-4, atol=1e-7)
    assert np.mean(gpu["last_count"] == whole["last_count"]) >= 0.9999
    ora = H.run_oracle(cfg.camera, s, inst)
    H.assert_bit_exact(gpu, ora)
    _image_checks(gpu, ora)


def test_a_frame_that_outgrows_the_pair_buffers_is_the_oracle_s():
    """The duplication runs behind the pair-count readback at the buffers'
    present size; a frame with more pairs runs it again after they grow."""
    print("pair buffers grown between frames", flush=True)
    r = raster.TileRasterizer()
    cfg = scenes.workload("plain-fwd-65k-640x480")
    s = raster.RenderSettings(width=cfg.width, height=cfg.height)
    small = scenes.workload("plain-fwd-65k-640x480", count=500)
    H.run_gpu(small.camera, s, [(small.scene, np.eye(4), raster.SplatEdit())], r)
    inst = [(cfg.scene, np.eye(4), raster.SplatEdit())]
    gpu = H.run_gpu(cfg.camera, s, inst, r)
    ora = H.run_oracle(cfg.camera, s, inst)
    assert ora["pairs"] > 20 * 500
    H.assert_bit_exact(gpu, ora)
    _image_checks(gpu, ora)


def test_walk_steps_are_counted_by_the_same_blend(rast):
    """LRT_DEBUG_WALK_STEPS runs the last frame's blend again and counts its
    quadrant waves' list steps: positive, no more than every quadrant walking its
    whole tile (each list fill padded to a group of four), the same on each call, and
    leaving the frame's images as they were."""
    cfg = scenes.workload("scene-fwd-3M-1080p", count=100_000)
    s = raster.RenderSettings(width=640, height=360, dilation=cfg.dilation, filter3d=cfg.filter3d)
    inst = [raster.SplatInstance(cfg.scene.to_device())]
    t = raster.RenderTargets(640, 360)
    rast.render(cfg.camera, inst, s, t)
    before = t.colour.clone()
    steps = rast.debug(_abi.DEBUG_WALK_STEPS)
    assert steps.size == 1
    ranges = rast.debug(_abi.DEBUG_RANGES).reshape(-1, 2).astype(np.int64)
    pairs = ranges[:, 1] - ranges[:, 0]
    bound = int((4 * (pairs + 4 * (pairs // 128 + 1))).sum())   # each list fill padded to a group of 4
    print(f"walk steps {int(steps[0])}, bound {bound}")
    assert 0 < int(steps[0]) <= bound
    assert int(rast.debug(_abi.DEBUG_WALK_STEPS)[0]) == int(steps[0])
    assert bool((t.colour == before).all())


def test_every_half_pattern_decodes_as_the_oracle_does(rast):
    """All 65,536 fp16 patterns (zeros, subnormals, infinities, quiet and
    signalling NaNs among them) as the 3,000-splat scene's harmonic
    coefficients, and a cycle of special halves as its base colours: the
    records, counts, pairs and ranges agree bit for bit with the oracle, whose
    decode is the exact bit-level one (the GPU converts in hardware and puts a
    NaN's quiet bit back from its payload)."""
    import dataclasses
    cloud = F.random_cloud(3000, 7).decode(15)   # degree 3: 23 words a splat, 138,000 halves
    halves = np.resize(np.arange(65536, dtype=np.uint32), cloud.sh.size * 2)
    sh = (halves[0::2] | (halves[1::2] << np.uint32(16))).reshape(cloud.sh.shape).astype(np.uint32)
    special = np.array([0x0000, 0x8000, 0x0001, 0x03FF, 0x8001, 0x7BFF, 0x7C00, 0xFC00, 0x7C01, 0x7E00, 0xFDFF,
                        0x3C00, 0x3555], np.uint32)
    shape = cloud.shape.copy()
    n = shape.shape[0]
    pick = special[np.arange(3 * n) % special.size].reshape(n, 3)
    shape[:, 2] = (shape[:, 2] & np.uint32(0xFFFF)) | (pick[:, 0] << np.uint32(16))   # z log-scale kept
    shape[:, 3] = pick[:, 1] | (pick[:, 2] << np.uint32(16))
    scene = dataclasses.replace(cloud, shape=shape, sh=sh)
    cam, _ = CAMERAS["perspective"]
    s = raster.RenderSettings(width=W, height=HGT)
    inst = [(scene, np.eye(4), raster.SplatEdit())]
    gpu = H.run_gpu(cam, s, inst, rast)
    ora = H.run_oracle(cam, s, inst)
    H.assert_bit_exact(gpu, ora)
