"""Runs one frame through the HIP library and through the oracle on the same
packed bytes, and returns both sides' buffers as host arrays."""
from __future__ import annotations

import numpy as np

import oracle_lib as O
from lucabrtrender_amd import _abi, raster


def oracle_camera(cam: raster.Camera) -> O.OrCamera:
    lens = cam.lens
    return O.camera(np.array(cam.camera_to_world), focal=lens.focal, orthographic=lens.orthographic,
                    haperture=lens.haperture, near_z=lens.near_z, far_z=lens.far_z,
                    window_translate=lens.window_translate, window_scale=lens.window_scale,
                    window_roll=lens.window_roll)


def oracle_edit(e: raster.SplatEdit) -> O.OrEdit:
    return O.edit(e.active, e.shape, e.mode, e.invert, e.centre, e.size, e.tint, e.saturation, e.brightness,
                  e.opacity, e.min_opacity, e.max_scale)


def oracle_settings(s: raster.RenderSettings) -> O.OrSettings:
    return O.settings(s.width, s.height, s.antialias, s.max_sh_degree, s.background, s.depth_mode,
                      s.depth_threshold, s.linearise, s.dilation, s.filter3d, s.tile_size)


def oracle_frames(cam, settings, instances):
    """instances: list of (HostScene, object_to_world 4x4, raster.SplatEdit)."""
    proj = O.projection_for(oracle_camera(cam), settings.width, settings.height)
    return O.frames_for(proj, oracle_settings(settings),
                        [(sc, np.asarray(o2w, dtype=np.float64), oracle_edit(ed)) for sc, o2w, ed in instances])


def run_gpu(cam, settings, instances, rast=None):
    import torch
    rast = rast or raster.TileRasterizer()
    dev = [sc.to_device() for sc, _, _ in instances]
    inst = [raster.SplatInstance(d, tuple(np.asarray(o2w, dtype=np.float64).reshape(-1)), ed)
            for d, (_, o2w, ed) in zip(dev, instances)]
    import dataclasses
    targets = raster.RenderTargets(settings.width, settings.height)
    stats = rast.render(cam, inst, dataclasses.replace(settings, keep_walk_state=True), targets)
    torch.cuda.synchronize()
    out = dict(stats=stats, radix_bits=settings.radix_bits,
               counts=rast.debug(_abi.DEBUG_COUNTS),
               touched=rast.debug(_abi.DEBUG_TOUCHED),
               offsets=rast.debug(_abi.DEBUG_OFFSETS),
               proj=rast.debug(_abi.DEBUG_PROJ).reshape(-1, 12),
               keys=rast.debug(_abi.DEBUG_KEYS),
               values=rast.debug(_abi.DEBUG_VALUES),
               ranges=rast.debug(_abi.DEBUG_RANGES).reshape(-1, 2),
               last_count=rast.debug(_abi.DEBUG_LAST_COUNT).reshape(settings.height, settings.width),
               final_t=rast.debug(_abi.DEBUG_FINAL_T).reshape(settings.height, settings.width),
               colour=targets.colour.cpu().numpy(),
               depth=targets.depth.cpu().numpy())
    # The same frame without the per-pixel walk state: the same images, bit for
    # bit (the walk only stops tracking its last contributor).
    fast = raster.RenderTargets(settings.width, settings.height)
    rast.render(cam, inst, dataclasses.replace(settings, keep_walk_state=False), fast)
    torch.cuda.synchronize()
    assert np.array_equal(fast.colour.cpu().numpy().view(np.uint32), out["colour"].view(np.uint32))
    assert np.array_equal(fast.depth.cpu().numpy().view(np.uint32), out["depth"].view(np.uint32))
    return out


def run_oracle(cam, settings, instances, threads=0):
    frames, clouds, n = oracle_frames(cam, settings, instances)
    out = O.raster(frames, clouds, n, threads=threads)
    out["frames"], out["clouds"], out["n"] = frames, clouds, n
    return out


def assert_bit_exact(gpu, ora):
    """Counts, projected records, sorted (tile, splat) pairs and ranges agree bit for bit."""
    n = ora["touched"].size
    assert gpu["counts"][0] == ora["splats"]
    assert gpu["counts"][1] == ora["visible"], (gpu["counts"], ora["visible"])
    assert gpu["counts"][2] == ora["pairs"], (gpu["counts"], ora["pairs"])
    np.testing.assert_array_equal(gpu["touched"][:n], ora["touched"])
    excl = np.concatenate([[0], np.cumsum(ora["touched"].astype(np.uint64))[:-1]]).astype(np.uint32)
    np.testing.assert_array_equal(gpu["offsets"][:n], excl)
    vis = ora["touched"] > 0
    np.testing.assert_array_equal(gpu["proj"][:n][vis].view(np.uint32), ora["proj"][vis].view(np.uint32))
    np.testing.assert_array_equal((gpu["keys"] >> np.uint64(32)).astype(np.uint32), ora["pair_tiles"])
    np.testing.assert_array_equal(gpu["values"], ora["pair_splats"])
    np.testing.assert_array_equal(gpu["ranges"], ora["ranges"])
    if gpu.get("stats") is not None and ora["pairs"] > 0:
        assert gpu["stats"].key_bits == expected_key_bits(ora), (gpu["stats"].key_bits, expected_key_bits(ora))
        assert gpu["stats"].sort_passes == expected_passes(gpu["stats"].key_bits, gpu.get("radix_bits", 0))


def expected_passes(key_bits: int, radix_bits: int = 0) -> int:
    """Sort passes: ceil(key bits / digit bits), the digit 9 bits by default
    where that saves a pass over 8."""
    radix = radix_bits or (9 if -(-key_bits // 9) < -(-key_bits // 8) else 8)
    return -(-key_bits // radix)


def bits_for(values: int) -> int:
    """The bits a key needs to tell `values` values apart (at least one)."""
    return max(1, (int(values) - 1).bit_length())


def expected_key_bits(ora) -> int:
    """The frame's sort key width: the tile id over the drawn slots' depth keys
    (bits(view z) >> 7) less their least, each field as wide as its range."""
    vis = ora["touched"] > 0
    depth_keys = (ora["proj"][vis][:, 2].view(np.uint32) & np.uint32(0x7FFFFFFF)) >> np.uint32(7)
    span = int(depth_keys.max()) - int(depth_keys.min()) + 1
    return bits_for(len(ora["ranges"])) + bits_for(span)


def oracle_style(st: raster.PointStyle) -> O.OrPointStyle:
    return O.point_style(st.size_mode, st.size, st.constant_colour, st.colour)


def oracle_composite_frames(cam, settings, instances, points):
    """instances as oracle_frames'; points: list of (positions float32 [N, 4],
    colours uint32 [N, 2] or None, object_to_world 4x4, raster.PointStyle), their
    slots after the splats'."""
    proj = O.projection_for(oracle_camera(cam), settings.width, settings.height)
    s = oracle_settings(settings)
    frames, clouds, n = O.frames_for(proj, s, [(sc, np.asarray(o2w, dtype=np.float64), oracle_edit(ed))
                                               for sc, o2w, ed in instances])
    base = sum(sc.count for sc, _, _ in instances)
    pframes, pts, npts = O.point_frames_for(proj, s, [(p, c, np.asarray(o2w, dtype=np.float64), oracle_style(st))
                                                      for p, c, o2w, st in points], base)
    return frames, clouds, n, pframes, pts, npts


def run_oracle_composite(cam, settings, instances, points, under_colour=None, under_depth=None, threads=0):
    frames, clouds, n, pframes, pts, npts = oracle_composite_frames(cam, settings, instances, points)
    return O.raster_composite(frames, clouds, n, pframes, pts, npts, under_colour, under_depth, threads=threads)


def run_reference_points(cam, settings, instances, points, threads=0):
    frames, clouds, n, pframes, pts, npts = oracle_composite_frames(cam, settings, instances, points)
    return O.reference_points(frames, clouds, n, pframes, pts, npts, threads=threads)


def run_gpu_composite(cam, settings, instances, points, under=None, rast=None):
    """The HIP composite frame: under is None or (colour [H, W, 4], depth [H, W])
    host arrays. Returns the images and the debug buffers."""
    import torch
    rast = rast or raster.TileRasterizer()
    dev = [sc.to_device() for sc, _, _ in instances]
    inst = [raster.SplatInstance(d, tuple(np.asarray(o2w, dtype=np.float64).reshape(-1)), ed)
            for d, (_, o2w, ed) in zip(dev, instances)]
    pinst = []
    for positions, colours, o2w, st in points:
        gp = raster.GpuPoints(torch.from_numpy(np.ascontiguousarray(positions)).cuda(),
                              torch.from_numpy(np.ascontiguousarray(colours).view(np.int32)).cuda()
                              if colours is not None else None)
        pinst.append(raster.PointInstance(gp, tuple(np.asarray(o2w, dtype=np.float64).reshape(-1)), st))
    targets = raster.RenderTargets(settings.width, settings.height)
    u = None
    if under is not None:
        u = raster.RenderTargets(settings.width, settings.height)
        u.colour.copy_(torch.from_numpy(np.ascontiguousarray(under[0], dtype=np.float32)))
        u.depth.copy_(torch.from_numpy(np.ascontiguousarray(under[1], dtype=np.float32)))
    stats = rast.render(cam, inst, settings, targets, points=pinst, under=u)
    torch.cuda.synchronize()
    return dict(stats=stats,
                counts=rast.debug(_abi.DEBUG_COUNTS),
                touched=rast.debug(_abi.DEBUG_TOUCHED),
                offsets=rast.debug(_abi.DEBUG_OFFSETS),
                proj=rast.debug(_abi.DEBUG_PROJ).reshape(-1, 12),
                keys=rast.debug(_abi.DEBUG_KEYS),
                values=rast.debug(_abi.DEBUG_VALUES),
                ranges=rast.debug(_abi.DEBUG_RANGES).reshape(-1, 2),
                colour=targets.colour.cpu().numpy(),
                depth=targets.depth.cpu().numpy())
