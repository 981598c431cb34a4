"""Batches of views of one FP32-parameter cloud: the differentiable forward and
the backward summed over the views.

- Each view of a batch is, bit for bit, the differentiable frame of that view
  alone, and the batch's gradients are, bit for bit, the single-view
  backwards' gradients added in view order (float32), under the grid blend,
  the persistent blend and the captured-graph replay.
- 4 views of mixed size x 1024 Gaussians against the FP64 oracles: each view's
  image against the FP64 brute force (p99 <= 1e-3, max <= 1e-2 absolute per
  channel), the summed gradients against the sum of the FP64 autodiff
  gradients of each view (relative error in norm < 1e-2 a group).
- The batched fwd+bwd workload at its full size (2,000,000 clustered
  Gaussians, SH degree 3, 64 cameras on a sphere of radius 12 at sizes drawn
  from 256x256 .. 1920x1080): every view's pair count and image equal its
  single-view frame's, and the batch's gradients equal the sum of the 64
  single-view backwards.
- A batch backward needs a batch: the single-view backward after a batch, the
  batch backward after a single frame and a wrong gradient count are refused.
"""
import numpy as np
import pytest
import torch

import harness as H
import oracle_lib as O
from lucabrtrender_amd import _abi, raster, scenes
from lucabrtrender_amd.autograd import rasterize_views
from lucabrtrender_amd._abi import LrtError
from oracle import diff_oracle as D

pytestmark = pytest.mark.gpu

GROUPS = ("means", "log_scales", "quats", "opacity_logits", "sh")
SIZES = [(128, 96), (96, 64), (160, 120), (64, 64)]


def _cams():
    eyes = [(0.0, 0.0, 0.0), (0.6, 0.3, 0.2), (-0.5, -0.2, 0.4), (0.2, 0.7, -0.5)]
    out = []
    for k, eye in enumerate(eyes):
        cam = raster.Camera.looking_at(eye, (0.0, 0.0, -4.0))
        cam.lens = raster.Lens(focal=scenes.focal_for_fov(60.0 - 5.0 * k))
        out.append(cam)
    return out


def _device(p):
    return {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in p.items()}


def _splats(dev, f3=None):
    return raster.DiffSplats(dev["means"], dev["log_scales"], dev["quats"], dev["opacity_logits"], dev["sh"],
                             filter3d=f3)


def _grads_like(dev):
    return {k: torch.empty_like(v) for k, v in dev.items()}


def _grad_colours(sizes, seed):
    gen = torch.Generator(device="cuda").manual_seed(seed)
    return [torch.randn((h, w, 4), generator=gen, device="cuda", dtype=torch.float32) for w, h in sizes]


def _view_settings(settings, size):
    return raster.RenderSettings(**{**settings.__dict__, "width": size[0], "height": size[1]})


def _batch_vs_singles(p, cams, sizes, settings, f3=None, seed=0):
    """(views, batch stats, batch grads, per-view single (colour, pairs), summed single grads)."""
    dev = _device(p)
    splats = _splats(dev, f3)
    rast = raster.TileRasterizer()
    views = [raster.View(c, raster.RenderTargets(w, h)) for c, (w, h) in zip(cams, sizes)]
    stats = rast.render_diff_views(views, splats, settings)
    gcols = _grad_colours(sizes, seed)
    grads = _grads_like(dev)
    rast.backward_views(gcols, grads)
    ranges = rast.debug(_abi.DEBUG_RANGES).reshape(-1, 2).astype(np.int64)
    torch.cuda.synchronize()
    single = raster.TileRasterizer()
    summed = None
    per_view = []
    tile_base = 0
    for v, (cam, (w, h)) in enumerate(zip(cams, sizes)):
        tiles = ((w + 15) // 16) * ((h + 15) // 16)
        batch_pairs = int((ranges[tile_base:tile_base + tiles, 1] - ranges[tile_base:tile_base + tiles, 0]).sum())
        tile_base += tiles
        t = raster.RenderTargets(w, h)
        st = single.render_diff(cam, splats, _view_settings(settings, (w, h)), t)
        g = _grads_like(dev)
        single.backward(gcols[v], g)
        assert torch.equal(t.colour, views[v].targets.colour), f"view {v}: colour differs from its single frame"
        assert torch.equal(t.depth, views[v].targets.depth), f"view {v}: depth differs from its single frame"
        assert batch_pairs == st.pairs, (v, batch_pairs, st.pairs)
        per_view.append(st.pairs)
        summed = g if summed is None else {k: summed[k] + g[k] for k in GROUPS}
    assert sum(per_view) == stats.pairs
    return views, stats, grads, summed, gcols


@pytest.mark.parametrize("persistent,graph", [(False, False), (True, False), (False, True)])
def test_each_view_is_its_single_frame_and_the_gradients_their_sum(persistent, graph):
    p = scenes.diff_params(1024, 51, degree=3, log_scale=(-3.0, 0.4))
    settings = raster.RenderSettings(persistent_blend=persistent, use_graph=graph, background=(0.1, 0.2, 0.05, 0.5))
    print(f"persistent {persistent} graph {graph}", flush=True)
    _, stats, grads, summed, _ = _batch_vs_singles(p, _cams(), SIZES, settings)
    assert stats.pairs > 1000
    for k in GROUPS:
        assert torch.equal(grads[k], summed[k]), k


def test_batch_against_the_fp64_oracles():
    p = scenes.diff_params(1024, 61, degree=3, log_scale=(-3.0, 0.4))
    n = p["means"].shape[0]
    f3 = np.full(n, 0.006, np.float32)
    settings = raster.RenderSettings(filter3d=True, linearise=True, dilation=0.1)
    cams = _cams()
    dev = _device(p)
    rast = raster.TileRasterizer()
    views = [raster.View(c, raster.RenderTargets(w, h)) for c, (w, h) in zip(cams, SIZES)]
    rast.render_diff_views(views, _splats(dev, torch.from_numpy(f3).cuda()), settings)
    gcols = _grad_colours(SIZES, 9)
    grads = _grads_like(dev)
    rast.backward_views(gcols, grads)
    torch.cuda.synchronize()
    ref = {k: np.zeros(v.shape, np.float64) for k, v in p.items()}
    for v, (cam, (w, h)) in enumerate(zip(cams, SIZES)):
        s = _view_settings(settings, (w, h))
        frame = O.diff_frame(O.projection_for(H.oracle_camera(cam), w, h), H.oracle_settings(s), n, 15)
        img = views[v].targets.colour.cpu().numpy().astype(np.float64)
        d = np.abs(img - O.diff_render_f64(frame, p, filter3d=f3)).reshape(-1)
        print(f"view {v} {w}x{h}: image vs FP64 brute force p99 {np.quantile(d, 0.99):.2e} max {d.max():.2e}")
        assert np.quantile(d, 0.99) <= 1e-3 and d.max() <= 1e-2
        g = D.gradients(frame, p, gcols[v].cpu().numpy().astype(np.float64), filter3d=f3)
        for k in GROUPS:
            ref[k] += g[k]
    for k in GROUPS:
        a, b = grads[k].cpu().numpy().astype(np.float64).reshape(-1), ref[k].reshape(-1)
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        print(f"  {k:15s} summed over 4 views, relative error vs FP64 autodiff {rel:.2e}")
        assert rel < 1e-2, (k, rel)


def test_full_size_batch_2m_splats_64_views():
    """The batched fwd+bwd workload at its full size (batch-fwd-bwd-2M-64)."""
    p = scenes.diff_params_clustered(2_000_000, 0, degree=3)
    views = scenes.sphere_views(64, 12.0, seed=4)
    cams = [c for c, _ in views]
    sizes = [sz for _, sz in views]
    f3 = scenes.filter3d_for(p["means"], cams[::4], 1920, 1080)
    settings = raster.RenderSettings(dilation=0.1, filter3d=True)
    print(f"2M splats, 64 views, {sum(w * h for w, h in sizes)} pixels", flush=True)
    _, stats, grads, summed, _ = _batch_vs_singles(p, cams, sizes, settings, f3=torch.from_numpy(f3).cuda(), seed=3)
    print(f"  {stats.pairs} pairs over the batch, {stats.visible} drawn slots")
    for k in GROUPS:
        assert torch.equal(grads[k], summed[k]), k
        assert bool(torch.isfinite(grads[k]).all()), k
    assert float(grads["means"].abs().sum()) > 0.0


def test_autograd_over_a_batch_matches_the_c_abi():
    p = scenes.diff_params(1024, 71, degree=3, log_scale=(-3.0, 0.4))
    settings = raster.RenderSettings()
    cams = _cams()
    rast = raster.TileRasterizer()
    dev = _device(p)
    views = [raster.View(c, raster.RenderTargets(w, h)) for c, (w, h) in zip(cams, SIZES)]
    rast.render_diff_views(views, _splats(dev), settings)
    gcols = _grad_colours(SIZES, 5)
    direct = _grads_like(dev)
    rast.backward_views(gcols, direct)
    leaves = {k: v.clone().requires_grad_(True) for k, v in dev.items()}
    colours = rasterize_views(leaves["means"], leaves["log_scales"], leaves["quats"], leaves["opacity_logits"],
                              leaves["sh"], rast, cams, SIZES, settings)
    for c, v in zip(colours, views):
        assert torch.equal(c, v.targets.colour)
    sum((c * g).sum() for c, g in zip(colours, gcols)).backward()
    for k in GROUPS:
        assert torch.equal(leaves[k].grad, direct[k]), k


def test_a_batch_backward_needs_a_batch():
    p = scenes.diff_params(256, 81, degree=3, log_scale=(-3.0, 0.4))
    dev = _device(p)
    splats = _splats(dev)
    rast = raster.TileRasterizer()
    cams = _cams()
    views = [raster.View(c, raster.RenderTargets(w, h)) for c, (w, h) in zip(cams, SIZES)]
    rast.render_diff_views(views, splats, raster.RenderSettings())
    grads = _grads_like(dev)
    gcols = _grad_colours(SIZES, 1)
    with pytest.raises(LrtError) as err:
        rast.backward_views(gcols[:3], grads)             # one gradient a view
    assert err.value.code == 3
    with pytest.raises(LrtError) as err:
        rast.backward(gcols[0], grads)                     # the single-view backward after a batch
    assert err.value.code == 3
    rast.backward_views(gcols, grads)                      # still kept after the refusals
    t = raster.RenderTargets(*SIZES[0])
    rast.render_diff(cams[0], splats, _view_settings(raster.RenderSettings(), SIZES[0]), t)
    with pytest.raises(LrtError) as err:
        rast.backward_views(gcols, grads)                  # the batch backward after a single frame
    assert err.value.code == 3
