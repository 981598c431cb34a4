#!/usr/bin/env python3
"""Forward frames/s of the splat tile rasteriser on one MI355X.

Workload: the 3,000,000-splat clustered scene at 1920x1080 (256 anisotropic
clusters, SH degree 3 in fp16, Mip-Splatting 3D filter on, 2D filter 0.1,
camera on an orbit of radius 15), synthetic and seeded, resident in HBM before
the timed region. A step is one whole frame through the C ABI: preprocess,
scan, duplication into 64-bit keys, onesweep sort, ranges, blend, including
the frame's own pair-count readback and final wait.

Prints one JSON line: frames/s, per-kernel GPU times from HIP events on the
rasteriser's stream with each stage's algorithmic bytes and fraction of HBM
peak, the dominant kernel's roofline, sort Gkeys/s, and the CPU baseline (the
oracle's restatement of the reference forward path on the host cores).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
FP32_PEAK_TFLOPS = 157.3       # MI355X vector FP32
METRIC = "fwd frames/s @3M splats 1080p; fwd+bwd ms/view; sort Gkeys/s; %HBM peak"
STAGES = ["project", "scan", "duplicate", "sort_histogram", "sort_passes", "ranges", "blend"]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def sort_pass_bytes(pairs, passes, key_bits):
    """HBM bytes of a sort's passes: each reads and writes its pairs. A key at
    most 32 bits wider than its first digit (ceil(key bits / passes) bits) leaves
    the first pass as u32 (12 bytes in, 8 out; 8 each way after), a wider one
    moves 12 bytes each way every pass."""
    if passes == 0:
        return 0
    if key_bits <= 32 + -(-key_bits // passes):
        return pairs * (12 + 8 + 16 * (passes - 1))
    return pairs * 24 * passes


def bits_for(values):
    """The bits a key needs to tell `values` values apart (at least one)."""
    return max(1, (max(values, 1) - 1).bit_length())


def defined_sort_bytes(pairs, tiles, radix_bits=8):
    """The sort's algorithmic bytes as the metric defines them: a stable onesweep
    of a 64-bit (tile, 24-bit depth key) key and a 32-bit value, P * (8 + 24 p)
    with p = ceil((tile bits + 24) / radix bits) passes (the histogram reads the
    8-byte keys once; each pass reads and writes 12 bytes a pair)."""
    p = -(-(bits_for(tiles) + 24) // radix_bits)
    return pairs * (8 + 24 * p), p


def stage_bytes(n, visible, pairs, tiles, pixels, sh_words, passes, key_bits):
    """Algorithmic HBM bytes of each kernel for one frame."""
    return {
        # every splat's packed record and 3D filter read, rect (8) and count (4)
        # written; the drawn ones' harmonics read and record (48) written
        "project": n * (32 + 4 + 12) + visible * (4 * sh_words + 48),
        # one u32 read and one written per splat
        "scan": n * 8,
        # the visible splats' record, rect, count and offset read; key + value written per pair
        "duplicate": visible * (16 + 16 + 8 + 4 + 4) + pairs * 12,
        # every key read once
        "sort_histogram": pairs * 8,
        # each pass reads and writes its pairs (8 or 12 bytes each way)
        "sort_passes": sort_pass_bytes(pairs, passes, key_bits),
        # keys read, two u32 a tile written; the ordering reads them again and writes a u32 a tile
        "ranges": pairs * 8 + tiles * 8 + tiles * 12,
        # each pair's slot and record read once a tile, the pixel out (colour 16 + depth 4; the
        # walk state is kept only on request)
        "blend": pairs * (4 + 48) + pixels * 20,
    }


KERNEL_OF_STAGE = {"project": "projectKernel", "scan": "scanKernel", "duplicate": "duplicateKernel",
                   "sort_histogram": "histogramKernel", "sort_passes": "onesweepKernel", "ranges": "rangesKernel",
                   "blend": "blendKernel"}


def measured_traffic_bytes(stage, passes):
    """HBM bytes one launch of the stage's kernel moved, from the committed
    rocprofv3 --pmc summary (32 x the DRAM read and write request counts in
    32-byte units), or None without one."""
    path = os.path.join(ROOT, "profiles", "traffic_scene-fwd-3M-1080p.txt")
    try:
        text = open(path).read()
    except OSError:
        return None
    for line in text.splitlines():
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 5 and cells[0].startswith(KERNEL_OF_STAGE[stage]):
            mb = float(cells[4].replace(",", ""))
            return int(mb * 1e6 * (passes if stage == "sort_passes" else 1))
    return None


# FP32 operations of one step of the blend's walk for one pixel (csrc/blend.hip,
# the splat-only body): the exponent's quadratic form (5 FMA), exp2, the 0.99
# clamp, T * alpha, the rgb and depth sums (4 FMA), T - weight. An FMA counts 2.
WALK_FLOPS = 5 * 2 + 1 + 1 + 1 + 4 * 2 + 1


def roofline_of(kernel, stages, passes):
    """The dominant kernel against the ceiling that bounds it: the blend against
    MI355X's FP32 VALU peak (its walk is VALU-bound; achieved = walk steps x 64
    pixels x WALK_FLOPS / the launch's time), any other kernel against HBM
    (achieved = algorithmic bytes / time). traffic: the kernel's measured HBM
    bytes a launch (rocprofv3 counters, the committed traffic summary)."""
    st = stages[kernel]
    traffic = measured_traffic_bytes(kernel, passes)
    if kernel == "blend":
        achieved = st["TFLOP/s"]
        return {"kernel": kernel, "bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / FP32_PEAK_TFLOPS, 4), "traffic": traffic,
                "walk_steps": st["walk_steps"], "flops_per_step": 64 * WALK_FLOPS,
                "hbm": {"algorithmic_bytes": st["bytes"], "GB/s": st["GB/s"], "frac": st["hbm_frac"]}}
    return {"kernel": kernel, "bound": "hbm", "achieved": st["GB/s"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": st["hbm_frac"], "traffic": traffic}


def copy_bandwidth_gbs(torch):
    """Achieved device-to-device copy bandwidth (read + write bytes / time)."""
    n = 1 << 28   # 1 GiB of float32
    a = torch.empty(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    reps = 10
    for _ in range(reps):
        b.copy_(a)
    stop.record()
    torch.cuda.synchronize()
    ms = start.elapsed_time(stop) / reps
    del a, b
    return 2 * n * 4 / (ms * 1e-3) / 1e9


def cpu_baseline(cfg, settings, seconds_budget=15.0):
    """The oracle's forward frame on the host cores, a bounded number of frames."""
    import ctypes as C

    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import harness as H
    import oracle_lib as O

    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    # Every core this process may run on, up to the host share its launcher
    # grants (OMP_NUM_THREADS when set).
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    threads = max(1, min(cores, share) if share > 0 else cores)
    from lucabrtrender_amd import raster
    frames, clouds, n = H.oracle_frames(cfg.camera, settings, [(cfg.scene, np.eye(4), raster.SplatEdit())])
    f0 = frames[0]
    total = cfg.scene.count
    proj = np.zeros((total, 12), np.float32)
    touched = np.zeros(total, np.uint32)
    keys = np.zeros(total, np.uint32)
    tiles = f0.tiles_x * f0.tiles_y
    ranges = np.zeros(tiles * 2, np.uint32)
    colour = np.zeros((f0.height, f0.width, 4), np.float32)
    depth = np.zeros((f0.height, f0.width), np.float32)
    stats = O.OrStats()

    def frame():
        return O.lib.or_raster(frames, clouds, n, O.ptr(proj), O.ptr(touched), O.ptr(keys), None, None, 0,
                               O.ptr(ranges), O.ptr(colour), O.ptr(depth), None, None, C.byref(stats), threads)

    frame()   # warm-up
    done, start = 0, time.perf_counter()
    while True:
        frame()
        done += 1
        elapsed = time.perf_counter() - start
        if elapsed >= seconds_budget or done >= 200:
            break
    model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"value": done / elapsed, "unit": "frames/s", "cores": threads, "kind": "port",
            "host_cores": os.cpu_count(), "cores_available": cores,
            "sample": f"{done} whole frames of the same 3M-splat 1080p scene ({elapsed:.1f} s), the oracle's C "
                      f"restatement of the reference forward path (project, 24-bit depth sort, emit, tile "
                      f"sort, ranges, blend) on {threads} OpenMP threads, {model}"}


def fwd_bwd_leg(torch, raster, scenes, steps):
    """fwd-bwd-1M-720p: 1,000,000 FP32-parameter splats of the plain cube with SH degree 3,
    1280x720; loss = per-pixel L2 against the frame of the perturbed scene. A view
    is one differentiable frame, the loss gradient and the backward."""
    params = scenes.diff_params(1_000_000, 0, degree=3)
    target_p = scenes.perturbed(params, seed=1)
    cam = raster.Camera.looking_at((0.0, 0.0, 0.0), (0.0, 0.0, -1.0))
    cam.lens = raster.Lens(focal=scenes.focal_for_fov(60.0))
    settings = raster.RenderSettings(width=1280, height=720)
    rast = raster.TileRasterizer()

    def device(p):
        d = {k: torch.from_numpy(v).cuda() for k, v in p.items()}
        return d, raster.DiffSplats(d["means"], d["log_scales"], d["quats"], d["opacity_logits"], d["sh"])

    _, tsplats = device(target_p)
    targets = raster.RenderTargets(1280, 720)
    rast.render_diff(cam, tsplats, settings, targets)
    target = targets.colour.clone()
    dev, splats = device(params)
    grads = {k: torch.empty_like(v) for k, v in dev.items()}

    def view():
        stats = rast.render_diff(cam, splats, settings, targets)
        grad = 2.0 * (targets.colour - target)
        return stats, rast.backward(grad, grads)

    for _ in range(3):
        view()
    torch.cuda.synchronize()
    start = time.perf_counter()
    bwd = [0.0, 0.0]
    for _ in range(steps):
        stats, ms = view()
        bwd[0] += ms[0]
        bwd[1] += ms[1]
    torch.cuda.synchronize()
    per_view = (time.perf_counter() - start) / steps
    start = time.perf_counter()
    for _ in range(steps):
        rast.render_diff(cam, splats, settings, targets)
    torch.cuda.synchronize()
    fwd = (time.perf_counter() - start) / steps
    # The same scene packed as the plain forward draws it (fp16 shape and
    # harmonics, linear opacity, base colour 0.5 + C0 dc), through that forward.
    import numpy as np
    n = params["means"].shape[0]
    fields = np.concatenate([params["means"], 1.0 / (1.0 + np.exp(-params["opacity_logits"][:, None])),
                             params["log_scales"], params["quats"], 0.5 + scenes.SH_C0 * params["sh"][:, 0, :],
                             params["sh"][:, 1:, :].reshape(n, -1)], axis=1).astype(np.float32)
    packed = scenes.pack_splats(fields, 15).to_device()
    inst = [raster.SplatInstance(packed)]
    plain_t = raster.RenderTargets(1280, 720)
    for _ in range(3):
        plain_stats = rast.render(cam, inst, settings, plain_t)
    torch.cuda.synchronize()
    start = time.perf_counter()
    for _ in range(steps):
        rast.render(cam, inst, settings, plain_t)
    torch.cuda.synchronize()
    plain = (time.perf_counter() - start) / steps
    return {"workload": "fwd-bwd-1M-720p", "splats": 1_000_000, "visible": stats.visible, "pairs": stats.pairs,
            "ms_per_view": round(per_view * 1e3, 4), "forward_ms": round(fwd * 1e3, 4),
            "blend_backward_ms": round(bwd[0] / steps, 4), "project_backward_ms": round(bwd[1] / steps, 4),
            "fwd_bwd_over_fwd": round(per_view / fwd, 3),
            "plain_forward_ms": round(plain * 1e3, 4), "plain_forward_pairs": plain_stats.pairs,
            "fwd_bwd_over_plain_fwd": round(per_view / plain, 3),
            "data": "synthetic: seeded FP32 parameters, target from means+N(0,0.01), logits+N(0,0.2), SH+N(0,0.05)"}


def batch_fwd_bwd_leg(torch, raster, scenes, reps):
    """batch-fwd-bwd-2M-64: one 2,000,000-Gaussian clustered scene of FP32 parameters
    (SH degree 3, 3D filter on, 2D filter 0.1), 64 cameras on a sphere of radius 12
    at sizes drawn from {256x256, 512x384, 960x540, 1280x720, 1920x1080}. Loss:
    per-pixel L2 against each view's frame of the perturbed scene. A step is the
    batched differentiable forward of all 64 views, the loss gradients and the
    batched backward (gradients summed over the views); against it the plain
    sequence of 64 single-view forward + backward steps."""
    params = scenes.diff_params_clustered(2_000_000, 0, degree=3)
    target_p = scenes.perturbed(params, seed=1)
    views_spec = scenes.sphere_views(64, 12.0, seed=4)
    cams = [c for c, _ in views_spec]
    f3 = torch.from_numpy(scenes.filter3d_for(params["means"], cams[::4], 1920, 1080)).cuda()
    settings = raster.RenderSettings(dilation=0.1, filter3d=True)
    rast = raster.TileRasterizer()

    def device(p):
        d = {k: torch.from_numpy(v).cuda() for k, v in p.items()}
        return d, raster.DiffSplats(d["means"], d["log_scales"], d["quats"], d["opacity_logits"], d["sh"],
                                    filter3d=f3)

    views = [raster.View(c, raster.RenderTargets(w, h)) for c, (w, h) in views_spec]
    _, tsplats = device(target_p)
    rast.render_diff_views(views, tsplats, settings)
    targets = [v.targets.colour.clone() for v in views]
    dev, splats = device(params)
    grads = {k: torch.empty_like(v) for k, v in dev.items()}

    def step():
        st = rast.render_diff_views(views, splats, settings)
        gcols = [2.0 * (v.targets.colour - t) for v, t in zip(views, targets)]
        return st, rast.backward_views(gcols, grads)

    for _ in range(2):
        step()
    torch.cuda.synchronize()
    start = time.perf_counter()
    bwd = [0.0, 0.0]
    for _ in range(reps):
        stats, ms = step()
        bwd[0] += ms[0]
        bwd[1] += ms[1]
    torch.cuda.synchronize()
    batch = (time.perf_counter() - start) / reps
    start = time.perf_counter()
    for _ in range(reps):
        rast.render_diff_views(views, splats, settings)
    torch.cuda.synchronize()
    batch_fwd = (time.perf_counter() - start) / reps
    # The same views one at a time: forward, loss gradient, backward.
    single = [raster.RenderSettings(**{**settings.__dict__, "width": v.targets.width, "height": v.targets.height})
              for v in views]
    one = {k: torch.empty_like(v) for k, v in dev.items()}

    def sequence():
        for v, s_, t in zip(views, single, targets):
            rast.render_diff(v.camera, splats, s_, v.targets)
            rast.backward(2.0 * (v.targets.colour - t), one)

    sequence()
    torch.cuda.synchronize()
    start = time.perf_counter()
    for _ in range(reps):
        sequence()
    torch.cuda.synchronize()
    seq = (time.perf_counter() - start) / reps
    n = len(views)
    return {"workload": "batch-fwd-bwd-2M-64", "splats": 2_000_000, "views": n,
            "pixels": sum(v.targets.width * v.targets.height for v in views), "pairs": stats.pairs,
            "batch_fwd_bwd_ms": round(batch * 1e3, 3), "fwd_bwd_ms_per_view": round(batch * 1e3 / n, 4),
            "batch_fwd_ms": round(batch_fwd * 1e3, 3), "fwd_ms_per_view": round(batch_fwd * 1e3 / n, 4),
            "fwd_bwd_over_fwd": round(batch / batch_fwd, 3),
            "blend_backward_ms": round(bwd[0] / reps, 3), "project_backward_ms": round(bwd[1] / reps, 3),
            "sequence_fwd_bwd_ms": round(seq * 1e3, 3), "speedup_over_sequence": round(seq / batch, 3),
            "data": "synthetic: seeded clustered FP32 parameters; target from means+N(0,0.01), logits+N(0,0.2), "
                    "SH+N(0,0.05)"}


def _batch_against_sequence(torch, raster, rast, views, inst, settings, reps):
    """One autotuned batch of the views against the plain sequence of single-view
    frames of the same views; the batch's eager stage times."""
    import numpy as np
    from lucabrtrender_amd import _abi
    tuned, _ = rast.autotune_views(views, inst, settings)
    rast.render_views(views, inst, tuned)
    torch.cuda.synchronize()
    start = time.perf_counter()
    for _ in range(reps):
        stats = rast.render_views(views, inst, tuned)
    torch.cuda.synchronize()
    batch_s = (time.perf_counter() - start) / reps
    single = [raster.RenderSettings(**{**settings.__dict__, "width": v.targets.width, "height": v.targets.height})
              for v in views]
    for v, s in zip(views, single):
        rast.render(v.camera, inst, s, v.targets)
    torch.cuda.synchronize()
    start = time.perf_counter()
    for _ in range(reps):
        for v, s in zip(views, single):
            rast.render(v.camera, inst, s, v.targets)
    torch.cuda.synchronize()
    seq_s = (time.perf_counter() - start) / reps
    staged = raster.RenderSettings(**{**tuned.__dict__, "use_graph": False, "time_stages": True})
    acc = np.zeros(8)
    for _ in range(reps):
        rast.render_views(views, inst, staged)
        acc += rast.debug(_abi.DEBUG_KERNEL_MS)
    n = len(views)
    return {"views": n, "pixels": sum(v.targets.width * v.targets.height for v in views), "pairs": stats.pairs,
            "batch_ms": round(batch_s * 1e3, 3), "views_per_s_batched": round(n / batch_s, 1),
            "sequence_ms": round(seq_s * 1e3, 3), "views_per_s_single": round(n / seq_s, 1),
            "speedup": round(seq_s / batch_s, 3),
            "stage_ms_eager": {name: round(float(acc[k] / reps), 3) for k, name in enumerate(STAGES)},
            "tuned": {"blend_batch": tuned.blend_batch, "persistent_blend": tuned.persistent_blend,
                      "use_graph": tuned.use_graph, "tile_size": tuned.tile_size}}


def views_leg(torch, raster, scenes, reps):
    """views-fwd-2M-64: one 2,000,000-splat clustered scene, 64 cameras on a sphere of radius
    12 looking at its centre, each view's size drawn from {256x256, 512x384, 960x540,
    1280x720, 1920x1080}. The batch (one launch sequence, autotuned) against the plain
    sequence of 64 single-view frames."""
    cfg = scenes.workload("scene-fwd-3M-1080p", count=2_000_000)
    inst = [raster.SplatInstance(cfg.scene.to_device())]
    views = [raster.View(cam, raster.RenderTargets(w, h)) for cam, (w, h) in scenes.sphere_views(64, 12.0, seed=4)]
    settings = raster.RenderSettings(dilation=cfg.dilation, filter3d=cfg.filter3d)
    out = _batch_against_sequence(torch, raster, raster.TileRasterizer(), views, inst, settings, reps)
    return {"workload": "views-fwd-2M-64", "splats": 2_000_000, **out}


def views_1080p_leg(torch, raster, scenes, cfg, dev, reps):
    """The batching target's own workload: the benchmark's 3,000,000-splat scene seen
    by 64 cameras at 1920x1080 on its radius-15 orbit, every 5.625 degrees, as one
    batch against the plain sequence of 64 single-view frames."""
    inst = [raster.SplatInstance(dev)]
    views = [raster.View(scenes.orbit_camera(15.0, 360.0 * k / 64.0), raster.RenderTargets(cfg.width, cfg.height))
             for k in range(64)]
    settings = raster.RenderSettings(dilation=cfg.dilation, filter3d=cfg.filter3d)
    out = _batch_against_sequence(torch, raster, raster.TileRasterizer(), views, inst, settings, reps)
    return {"workload": "views-fwd-3M-1080p-64", "splats": cfg.scene.count, **out}


def composite_leg(torch, raster, _abi, cfg, dev, settings, reps):
    """Composite frames on the benchmark scene (3,000,000 splats at 1920x1080):
    1,000,000 points near its splat centres (every third centre moved by a
    seeded N(0, 0.05) offset a coordinate) drawn as opaque 2-pixel discs with
    seeded colours, (a) in one sort with the splats, (b) alone, as a layer, and
    (c) the splats over that layer given as the under-layer. Frames/s, pairs
    and the blend's GPU time of each, against the splat-only frame; and the
    share of pixels where (c) and (a) differ by more than 1/255, which the two
    routes' equal arithmetic leaves to pairs whose 24-bit depth keys tie."""
    import numpy as np
    n = 1_000_000
    rng = np.random.default_rng(5)
    xyz = dev.positions[::3][:n, :3].cpu().numpy() + rng.normal(0.0, 0.05, (n, 3)).astype(np.float32)
    rgb = rng.random((n, 3), dtype=np.float32)
    positions, colours = raster.pack_points(xyz, rgb)
    points = raster.GpuPoints(torch.from_numpy(positions).cuda(), torch.from_numpy(colours.view(np.int32)).cuda())
    pinst = [raster.PointInstance(points, style=raster.PointStyle(size_mode=raster.PointStyle.PIXELS, size=2.0))]
    inst = [raster.SplatInstance(dev)]
    rast = raster.TileRasterizer()
    proj = raster.projection_for(cfg.camera, cfg.width, cfg.height)
    mixed_t = raster.RenderTargets(cfg.width, cfg.height)
    layer_t = raster.RenderTargets(cfg.width, cfg.height)
    over_t = raster.RenderTargets(cfg.width, cfg.height)
    plain_t = raster.RenderTargets(cfg.width, cfg.height)
    runs = {"splats_only": lambda s: rast.render(proj, inst, s, plain_t),
            "mixed": lambda s: rast.render(proj, inst, s, mixed_t, points=pinst),
            "discs_layer": lambda s: rast.render(proj, [], s, layer_t, points=pinst),
            "over_layer": lambda s: rast.render(proj, inst, s, over_t, under=layer_t)}
    timed = raster.RenderSettings(**{**settings.__dict__, "time_stages": True})
    out = {"workload": "scene-fwd-3M-1080p+1M-discs", "discs": n}
    for name, run in runs.items():
        for _ in range(2):
            stats = run(settings)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            stats = run(settings)
        torch.cuda.synchronize()
        fps = reps / (time.perf_counter() - t0)
        blend = 0.0
        for _ in range(reps):
            run(timed)
            blend += float(rast.debug(_abi.DEBUG_KERNEL_MS)[6])
        out[name] = {"fps": round(fps, 1), "pairs": stats.pairs, "visible": stats.visible,
                     "blend_ms": round(blend / reps, 4)}
    runs["mixed"](settings)
    runs["over_layer"](settings)
    torch.cuda.synchronize()
    diff = (mixed_t.colour - over_t.colour).abs().amax(dim=2)
    out["over_vs_mixed_pixels_over_1_255"] = round(float((diff > 1.0 / 255.0).float().mean()), 6)
    return out


def lod_leg(torch, raster, scenes, cfg, dev, settings, reps):
    """Levels of detail on the benchmark scene (3,000,000 splats): the build's
    wall time, then a far view at 1920x1080 (orbit radius 60) drawn whole and
    through cuts at 2, 4 and 8 pixels: the cut's wall time (its one wait for
    the counts included), the frame's, what was drawn, and the mean absolute
    colour difference from the whole cloud's frame. Both frames are drawn
    without the 3D filter, which merged Gaussians do not carry."""
    settings = raster.RenderSettings(**{**settings.__dict__, "filter3d": False})
    rast = raster.TileRasterizer()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    lod = rast.lod_build(dev)
    build_ms = (time.perf_counter() - t0) * 1e3
    cam = scenes.orbit_camera(60.0, 0.0)
    proj = raster.projection_for(cam, cfg.width, cfg.height)
    full_t = raster.RenderTargets(cfg.width, cfg.height)
    cut_t = raster.RenderTargets(cfg.width, cfg.height)
    inst = [raster.SplatInstance(dev)]
    for _ in range(2):
        rast.render(proj, inst, settings, full_t)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        rast.render(proj, inst, settings, full_t)
    torch.cuda.synchronize()
    out = {"workload": "scene-lod-3M-1080p-far", "splats": lod.count, "merged": lod.merged,
           "levels": [lv["level"] for lv in lod.levels], "build_ms": round(build_ms, 2),
           "off": {"render_ms": round((time.perf_counter() - t0) * 1e3 / reps, 4)}}
    for threshold in (2.0, 4.0, 8.0):
        linst = [raster.LodInstance(lod)]
        for _ in range(2):
            drawn, stats = rast.lod_select(proj, linst, threshold, settings)
            rast.render(proj, drawn, settings, cut_t)
        torch.cuda.synchronize()
        cut_ms = frame_ms = 0.0
        for _ in range(reps):
            t0 = time.perf_counter()
            drawn, stats = rast.lod_select(proj, linst, threshold, settings)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            rast.render(proj, drawn, settings, cut_t)
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            cut_ms += (t1 - t0) * 1e3
            frame_ms += (t2 - t1) * 1e3
        mad = float((cut_t.colour - full_t.colour).abs().mean())
        out[f"{threshold:g}px"] = {"drawn": stats[0].splats + stats[0].merged, "splats": stats[0].splats,
                                   "merged": stats[0].merged, "cut_ms": round(cut_ms / reps, 4),
                                   "render_ms": round(frame_ms / reps, 4), "mean_abs": round(mad, 6)}
    lod.close()
    return out


def sort_stress_leg(torch, raster, scenes, _abi, reps):
    """sort-stress-6M-2160p: 6,000,000 splats at 3840x2160, 80% of the plain cube and 20% large splats
    near the camera; the frame's stages timed, the sort's keys/s and HBM fraction."""
    import numpy as np
    cfg = scenes.workload("sort-stress-6M-2160p")
    dev = cfg.scene.to_device()
    settings = raster.RenderSettings(width=cfg.width, height=cfg.height, time_stages=True)
    targets = raster.RenderTargets(cfg.width, cfg.height)
    rast = raster.TileRasterizer()
    inst = [raster.SplatInstance(dev)]
    rast.render(cfg.camera, inst, settings, targets)
    acc = np.zeros(8)
    start = time.perf_counter()
    for _ in range(reps):
        st = rast.render(cfg.camera, inst, settings, targets)
        acc += rast.debug(_abi.DEBUG_KERNEL_MS)
    wall = (time.perf_counter() - start) / reps
    ms = acc / reps
    n, visible, pairs, tiles = (int(v) for v in rast.debug(_abi.DEBUG_COUNTS))
    passes = st.sort_passes
    sort_ms = float(ms[3] + ms[4])
    # The metric's algorithmic bytes (64-bit keys with the full 24-bit depth key
    # every pass), and the bytes this sort moves (keys u32 after the first pass,
    # the depth field narrowed to the drawn keys' span).
    defined, defined_passes = defined_sort_bytes(pairs, tiles)
    pass_bytes = sort_pass_bytes(pairs, passes, st.key_bits)
    moved = pairs * 8 + pass_bytes   # the histogram reads every key once
    # The sort at each digit width the library offers (the default above).
    by_radix = {}
    for radix in (8, 9, 10):
        rs = raster.RenderSettings(**{**settings.__dict__, "radix_bits": radix})
        acc_r = np.zeros(8)
        for _ in range(reps):
            rst = rast.render(cfg.camera, inst, rs, targets)
            acc_r += rast.debug(_abi.DEBUG_KERNEL_MS)
        by_radix[str(radix)] = {"passes": rst.sort_passes, "sort_ms": round(float(acc_r[3] + acc_r[4]) / reps, 4)}
    return {"workload": "sort-stress-6M-2160p", "splats": n, "visible": visible, "pairs": pairs,
            "sort_passes": passes, "key_bits": st.key_bits, "sort_ms": round(sort_ms, 4),
            "sort_gkeys_per_s": round(pairs / (sort_ms * 1e-3) / 1e9, 3),
            "sort_bytes": int(defined), "sort_bytes_passes": defined_passes,
            "sort_GB/s": round(defined / (sort_ms * 1e-3) / 1e9, 1),
            "sort_hbm_frac": round(defined / (sort_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "moved_bytes": int(moved), "moved_GB/s": round(moved / (sort_ms * 1e-3) / 1e9, 1),
            "moved_hbm_frac": round(moved / (sort_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "passes_ms": round(float(ms[4]), 4),
            "passes_moved_hbm_frac": round(pass_bytes / (float(ms[4]) * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "sort_ms_by_radix": by_radix,
            "duplicate_ms": round(float(ms[2]), 4), "blend_ms": round(float(ms[6]), 4),
            "gpu_ms_per_frame": round(float(ms[7]), 4), "wall_ms_per_frame": round(wall * 1e3, 4)}


def load_leg(torch, raster, reps, cpu_sample):
    """The load path: 3,000,000 3DGS-PLY records at degree 3 (62 floats, 248 B a
    record) decoded into packed splats on the GPU. The decode's kernels are
    timed with HIP events (records already on the device); the PCIe upload is
    reported beside it. cpu_sample records go through the oracle's decode on one
    host core (0 skips it)."""
    import ctypes as C

    import numpy as np
    n, floats = 3_000_000, 62
    rng = np.random.default_rng(11)
    rec = np.empty((n, floats), np.float32)
    rec[:, 0:3] = rng.uniform(-10, 10, (n, 3))
    rec[:, 3:51] = rng.normal(0.0, 0.3, (n, 48))              # dc + 45 rest, colour-outer
    rec[:, 51] = rng.normal(0.0, 2.0, n)                      # opacity logits: some below 1/255
    rec[:, 52:55] = rng.normal(-5.0, 0.8, (n, 3))             # log scales
    rec[:, 55:59] = rng.normal(0.0, 1.0, (n, 4))              # rotation w x y z
    rec[:, 59:62] = 0.0                                       # normals, ignored
    enc = dict(floats_per_record=floats, x=0, y=1, z=2, dc0=3, dc1=4, dc2=5, rest_base=6, rest_per_colour=15,
               file_per_colour=15, rest_colour_outer=1, opacity=51, scale0=52, scale1=53, scale2=54, rot_w=55,
               rot_x=56, rot_y=57, rot_z=58)
    raw = raster.raw_splats(rec, enc)
    rast = raster.TileRasterizer()
    rast.decode_cloud(raw, 3)   # warm-up: buffers sized
    up = dec = 0.0
    for _ in range(reps):
        _, info = rast.decode_cloud(raw, 3)
        up += info.upload_ms
        dec += info.decode_ms
    up, dec = up / reps, dec / reps
    kept = info.count
    # Algorithmic bytes: the validate pass's eight fields, the decode's whole
    # record, the valid flags and places (scan), the packed splat out.
    algo = n * (8 * 4 + floats * 4 + 4 * 4) + kept * (32 + 4 * info.sh_words)
    out = {"workload": "decode-3M-ply-sh3", "records": n, "kept": kept, "record_bytes": floats * 4,
           "decode_ms": round(dec, 4), "upload_ms": round(up, 4), "bytes": int(algo),
           "GB/s": round(algo / (dec * 1e-3) / 1e9, 1), "hbm_frac": round(algo / (dec * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
           "records_per_s": round(n / (dec * 1e-3), 1),
           "upload_GB/s": round(n * floats * 4 / (up * 1e-3) / 1e9, 1)}
    if cpu_sample > 0:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as O
        k = min(cpu_sample, n)
        e = O.OrEncoding(**{f: v for f, v in enc.items() if f != "rest_per_colour"})
        sample = np.ascontiguousarray(rec[:k])
        pos = np.zeros((k, 4), np.float32)
        shape = np.zeros((k, 4), np.uint32)
        sh = np.zeros((k, 23), np.uint32)
        start = time.perf_counter()
        O.lib.or_decode_records(O.ptr(sample), k, C.byref(e), 15, O.ptr(pos), O.ptr(shape), O.ptr(sh), 23)
        sec = time.perf_counter() - start
        out["cpu_baseline"] = {"value": round(k / sec, 1), "unit": "records/s", "cores": 1, "kind": "port",
                               "sample": f"the oracle's validate + decode of the first {k} records"}
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--count", type=int, default=3_000_000, help="splats (the workload is 3,000,000)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="CPU-baseline budget; 0 skips it")
    ap.add_argument("--blend-batch", type=int, default=256, help="records in each quadrant's LDS list: 64, 128 or 256")
    ap.add_argument("--tile-size", type=int, default=16, help="pixels a tile's side: 8, 16 (the reference's) or 32")
    ap.add_argument("--sweeps", type=int, default=1, help="1: the blend-setting and tile-size comparisons; 0 skips them")
    ap.add_argument("--image-pixels", type=int, default=65536,
                    help="pixels of the timed frame checked against the FP64 brute force; 0 skips it")
    ap.add_argument("--fwd-bwd-steps", type=int, default=20, help="fwd+bwd views (1M splats, 720p) timed; 0 skips it")
    ap.add_argument("--stress-reps", type=int, default=3, help="sort-stress frames (6M splats, 2160p) timed; 0 skips it")
    ap.add_argument("--views-reps", type=int, default=5, help="64-view batches (2M splats) timed; 0 skips it")
    ap.add_argument("--batch-bwd-reps", type=int, default=3,
                    help="batched fwd+bwd steps (2M splats, 64 views) timed; 0 skips it")
    ap.add_argument("--load-reps", type=int, default=3, help="3M-record decodes timed; 0 skips it")
    ap.add_argument("--lod-reps", type=int, default=5, help="LOD cuts and frames timed; 0 skips it")
    ap.add_argument("--composite-reps", type=int, default=5,
                    help="composite frames (1M discs with the 3M splats) timed; 0 skips it")
    ap.add_argument("--profile-frames", type=int, default=0,
                    help="only draw this many 3M-splat 1080p frames and one 256 MiB device copy (the counter "
                         "calibration), for rocprofv3 --pmc; prints no JSON")
    ap.add_argument("--profile-stress", type=int, default=0,
                    help="only draw this many sort-stress frames (6M splats, 2160p), for rocprofv3; prints no JSON")
    ap.add_argument("--profile-fwd-bwd", type=int, default=0,
                    help="only run this many fwd+bwd views (1M splats, 720p), for rocprofv3; prints no JSON")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build runs no multi-GPU path (asked for --gpus {args.gpus}); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from lucabrtrender_amd import _abi, raster, scenes

    if args.profile_stress > 0:
        cfg = scenes.workload("sort-stress-6M-2160p")
        inst = [raster.SplatInstance(cfg.scene.to_device())]
        settings = raster.RenderSettings(width=cfg.width, height=cfg.height)
        targets = raster.RenderTargets(cfg.width, cfg.height)
        rast = raster.TileRasterizer()
        for _ in range(args.profile_stress):
            stats = rast.render(cfg.camera, inst, settings, targets)
        torch.cuda.synchronize()
        log(f"{args.profile_stress} sort-stress frames drawn ({stats.pairs} pairs, {stats.sort_passes} passes)")
        return
    if args.profile_fwd_bwd > 0:
        fwd_bwd_leg(torch, raster, scenes, args.profile_fwd_bwd)
        log(f"{args.profile_fwd_bwd} fwd+bwd views run")
        return
    if args.profile_frames > 0:
        cfg = scenes.workload("scene-fwd-3M-1080p", count=args.count)
        settings = raster.RenderSettings(width=cfg.width, height=cfg.height, dilation=cfg.dilation,
                                         filter3d=cfg.filter3d, blend_batch=args.blend_batch)
        inst = [raster.SplatInstance(cfg.scene.to_device())]
        targets = raster.RenderTargets(cfg.width, cfg.height)
        rast = raster.TileRasterizer()
        for _ in range(args.profile_frames):
            rast.render(cfg.camera, inst, settings, targets)
        a = torch.empty(1 << 26, dtype=torch.float32, device="cuda")   # 256 MiB read, 256 MiB written
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        log(f"{args.profile_frames} frames and one 256 MiB copy drawn")
        return

    t0 = time.perf_counter()
    cfg = scenes.workload("scene-fwd-3M-1080p", count=args.count)
    log(f"scene: {cfg.scene.count} splats generated in {time.perf_counter() - t0:.1f} s")
    settings = raster.RenderSettings(width=cfg.width, height=cfg.height, dilation=cfg.dilation,
                                     filter3d=cfg.filter3d, blend_batch=args.blend_batch, tile_size=args.tile_size)
    dev = cfg.scene.to_device()
    inst = [raster.SplatInstance(dev)]
    targets = raster.RenderTargets(cfg.width, cfg.height)
    rast = raster.TileRasterizer()
    proj = raster.projection_for(cfg.camera, cfg.width, cfg.height)

    for _ in range(args.warmup):
        rast.render(proj, inst, settings, targets)
    torch.cuda.synchronize()
    start = time.perf_counter()
    for _ in range(args.steps):
        stats = rast.render(proj, inst, settings, targets)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - start
    fps = args.steps / elapsed
    log(f"{stats.splats} splats ({stats.visible} visible), {stats.pairs} pairs: {fps:.1f} frames/s")

    # Per-kernel GPU time: the same frames with HIP events between the kernels.
    timed = raster.RenderSettings(**{**settings.__dict__, "time_stages": True})
    reps = max(10, min(args.steps, 50))
    acc = np.zeros(8)
    for _ in range(reps):
        timed_stats = rast.render(proj, inst, timed, targets)
        acc += rast.debug(_abi.DEBUG_KERNEL_MS)
    kernel_ms = acc / reps
    counts = rast.debug(_abi.DEBUG_COUNTS)
    n, visible, pairs, tiles = (int(v) for v in counts)
    passes = timed_stats.sort_passes
    sbytes = stage_bytes(n, visible, pairs, tiles, cfg.width * cfg.height, cfg.scene.sh_words, passes,
                         timed_stats.key_bits)
    stages = {}
    for k, name in enumerate(STAGES):
        ms = float(kernel_ms[k])
        gbs = sbytes[name] / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
        stages[name] = {"ms": round(ms, 4), "bytes": int(sbytes[name]), "GB/s": round(gbs, 1),
                        "hbm_frac": round(gbs / HBM_PEAK_GBS, 4)}
    dominant = max(STAGES, key=lambda s: stages[s]["ms"])
    sort_ms = stages["sort_histogram"]["ms"] + stages["sort_passes"]["ms"]
    # The blend's walk steps (list records walked, each for a wave's 64 pixels),
    # counted by the same blend run again.
    walk_steps = int(rast.debug(_abi.DEBUG_WALK_STEPS)[0])
    stages["blend"]["walk_steps"] = walk_steps
    stages["blend"]["TFLOP/s"] = round(walk_steps * 64 * WALK_FLOPS / (stages["blend"]["ms"] * 1e-3) / 1e12, 3)
    roofline = roofline_of(dominant, stages, passes)
    copy_gbs = copy_bandwidth_gbs(torch)
    # The blend at each batch depth and launch order, split blends and the frame
    # at each tile size, for the record (--sweeps 0 skips them).
    variants, tile_sizes = {}, {}
    if args.sweeps:
        for batch in (64, 128, 256):
            for order in (0, 1):
                vs = raster.RenderSettings(**{**timed.__dict__, "blend_batch": batch, "tile_order": order})
                acc_v = np.zeros(8)
                for _ in range(10):
                    rast.render(proj, inst, vs, targets)
                    acc_v += rast.debug(_abi.DEBUG_KERNEL_MS)
                variants[f"batch{batch}_{'index_order' if order == 0 else 'heavy_first'}"] = round(float(acc_v[6] / 10), 4)
        # Heavy tiles split across waves: the blend at each segment length (0: never split).
        for segment in (0, 256, 512, 1024, 2048):
            vs = raster.RenderSettings(**{**timed.__dict__, "blend_segment": segment})
            acc_v = np.zeros(8)
            for _ in range(10):
                rast.render(proj, inst, vs, targets)
                acc_v += rast.debug(_abi.DEBUG_KERNEL_MS)
            variants[f"segment{segment}"] = round(float(acc_v[6] / 10), 4)
        # The frame at each tile size: the same images, different pair counts.
        for tile in (8, 16, 32):
            vs = raster.RenderSettings(**{**settings.__dict__, "tile_size": tile})
            for _ in range(3):
                rast.render(proj, inst, vs, targets)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            for _ in range(30):
                st = rast.render(proj, inst, vs, targets)
            torch.cuda.synchronize()
            tfps = 30 / (time.perf_counter() - t1)
            acc_v = np.zeros(8)
            for _ in range(10):
                rast.render(proj, inst, raster.RenderSettings(**{**vs.__dict__, "time_stages": True}), targets)
                acc_v += rast.debug(_abi.DEBUG_KERNEL_MS)
            tile_sizes[str(tile)] = {"frames_per_s": round(tfps, 1), "pairs": int(st.pairs),
                                     "stage_ms": {name: round(float(acc_v[k] / 10), 4) for k, name in enumerate(STAGES)}}
        log("tile sizes: " + ", ".join(f"{t}: {v['frames_per_s']} frames/s, {v['pairs']} pairs"
                                       for t, v in tile_sizes.items()))

    # The timed frame against the FP64 brute force over all splats on a seeded
    # pixel subset, in the rasteriser's order (24-bit depth key, slot) and in
    # the brute force's own (exact view z, index).
    image_vs_fp64 = None
    if args.image_pixels > 0:
        rast.render(proj, inst, settings, targets)
        npix = min(args.image_pixels, cfg.width * cfg.height)
        pix = np.random.default_rng(11).choice(cfg.width * cfg.height, npix, replace=False).astype(np.int32)
        ours = targets.colour.reshape(-1, 4).cpu().numpy().astype(np.float64)[pix]
        image_vs_fp64 = {"pixels": int(npix)}
        for keyed in (True, False):
            colour, _ = rast.reference_render_pixels(proj, inst, settings, torch.from_numpy(pix).cuda(), keyed=keyed)
            d = np.abs(ours - colour.cpu().numpy()).reshape(-1)
            image_vs_fp64["keyed_order" if keyed else "exact_order"] = {
                "p99": float(np.quantile(d, 0.99)), "max": float(d.max()), "over_1e-2": float(np.mean(d > 1e-2))}
        log(f"image vs FP64: {image_vs_fp64}")

    fwd_bwd = None
    if args.fwd_bwd_steps > 0:
        log("fwd+bwd (1M splats, 720p) ...")
        fwd_bwd = fwd_bwd_leg(torch, raster, scenes, args.fwd_bwd_steps)
    batch_fwd_bwd = None
    if args.batch_bwd_reps > 0:
        log("batched fwd+bwd (2M splats, 64 views) ...")
        batch_fwd_bwd = batch_fwd_bwd_leg(torch, raster, scenes, args.batch_bwd_reps)
    batched = None
    if args.views_reps > 0:
        log("batched views (2M splats, 64 views) ...")
        batched = views_leg(torch, raster, scenes, args.views_reps)
    batched_1080p = None
    if args.views_reps > 0:
        log("batched views (3M splats, 64 views at 1080p) ...")
        batched_1080p = views_1080p_leg(torch, raster, scenes, cfg, dev, max(2, args.views_reps // 2))
    stress = None
    if args.stress_reps > 0:
        log("sort stress (6M splats, 2160p) ...")
        stress = sort_stress_leg(torch, raster, scenes, _abi, args.stress_reps)

    composite = None
    if args.composite_reps > 0:
        log("composite frames (3M splats + 1M discs, 1080p) ...")
        composite = composite_leg(torch, raster, _abi, cfg, dev, settings, args.composite_reps)

    lod = None
    if args.lod_reps > 0:
        log("levels of detail (3M splats, far view at 1080p) ...")
        lod = lod_leg(torch, raster, scenes, cfg, dev, settings, args.lod_reps)

    load = None
    if args.load_reps > 0:
        log("load path (3M PLY records) ...")
        load = load_leg(torch, raster, args.load_reps, 300_000 if args.cpu_seconds > 0 else 0)

    cpu = None
    if args.cpu_seconds > 0:
        log("CPU baseline ...")
        cpu = cpu_baseline(cfg, settings, args.cpu_seconds)

    out = {
        "metric": METRIC,
        "value": round(fps, 2),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded 3M-splat clustered scene, packed fp16 shape/harmonics",
        "config": {"workload": "scene-fwd-3M-1080p", "splats": n, "visible": visible, "pairs": pairs,
                   "width": cfg.width, "height": cfg.height, "tiles": tiles, "sh_degree": 3, "filter3d": True,
                   "dilation": cfg.dilation, "sort_passes": passes, "key_bits": timed_stats.key_bits,
                   "tile_size": settings.tile_size},
        "stages": stages,
        "gpu_ms_per_frame": round(float(kernel_ms[7]), 4),
        "blend_ms_by_setting": variants,
        "tile_sizes": tile_sizes,
        "image_vs_fp64": image_vs_fp64,
        "sort_gkeys_per_s": round(pairs / (sort_ms * 1e-3) / 1e9, 3) if sort_ms > 0 else None,
        "copy_GB/s": round(copy_gbs, 1),
        "fwd_bwd": fwd_bwd,
        "batch_fwd_bwd": batch_fwd_bwd,
        "views": batched,
        "views_3m_1080p": batched_1080p,
        "sort_stress": stress,
        "load": load,
        "composite": composite,
        "lod": lod,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
