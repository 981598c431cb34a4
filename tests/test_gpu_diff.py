"""The differentiable frame on the GPU.

- Forward: the HIP image of an FP32-parameter cloud against the FP64 brute
  force (p99 <= 1e-3, max <= 1e-2 absolute per channel, linear FP32).
- Backward: the HIP gradients of sum(colour * G) for a random G against the
  FP64 autodiff oracle (oracle/diff_oracle.py, pinned on the CPU by
  tests/test_oracle_diff.py) and against FP64 central differences of the C
  brute force, with the relative error per parameter group printed and
  bounded.
- The torch.autograd wrapper gives the same gradients as the C ABI, and the
  context refuses a backward that no differentiable frame precedes.

The sub-scene is the finite-difference case of the 1M-splat 720p fwd+bwd
workload: 1,024 Gaussians of the plain cube with SH degree 3 and FP32 parameters, here at 128 x 96 with
log-scales N(-3, 0.4) so the footprints span a few pixels of the small frame.
"""
import numpy as np
import pytest
import torch

import oracle_lib as O
from lucabrtrender_amd import raster, scenes
from lucabrtrender_amd.autograd import rasterize
from lucabrtrender_amd._abi import LrtError
from oracle import diff_oracle as D

pytestmark = pytest.mark.gpu

W, HGT = 128, 96
GROUPS = ("means", "log_scales", "quats", "opacity_logits", "sh")


@pytest.fixture(scope="module")
def rast():
    return raster.TileRasterizer()


def _camera():
    cam = raster.Camera.looking_at((0.0, 0.0, 0.0), (0.0, 0.0, -1.0))
    cam.lens = raster.Lens(focal=scenes.focal_for_fov(60.0))
    return cam


def _oracle_frame(settings, count, rest):
    cam = O.camera(O.look_at((0.0, 0.0, 0.0), (0.0, 0.0, -1.0)), focal=scenes.focal_for_fov(60.0))
    proj = O.projection_for(cam, settings.width, settings.height)
    s = O.settings(settings.width, settings.height, settings.antialias, settings.max_sh_degree, settings.background,
                   settings.depth_mode, settings.depth_threshold, settings.linearise, settings.dilation,
                   settings.filter3d)
    return O.diff_frame(proj, s, count, rest)


def _device(p):
    return {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in p.items()}


def _gpu_grads(rast, p, settings, g, f3=None):
    dev = _device(p)
    f3d = torch.from_numpy(f3).cuda() if f3 is not None else None
    splats = raster.DiffSplats(dev["means"], dev["log_scales"], dev["quats"], dev["opacity_logits"], dev["sh"],
                               filter3d=f3d)
    targets = raster.RenderTargets(settings.width, settings.height)
    rast.render_diff(_camera(), splats, settings, targets)
    grads = {k: torch.empty_like(v) for k, v in dev.items()}
    gt = torch.from_numpy(np.ascontiguousarray(g, dtype=np.float32)).cuda()
    rast.backward(gt, grads)
    torch.cuda.synchronize()
    return targets.colour.cpu().numpy(), {k: v.cpu().numpy() for k, v in grads.items()}


CASES = [
    dict(linearise=True, filter3d=False, antialias=True, background=(0.0, 0.0, 0.0, 0.0)),
    dict(linearise=False, filter3d=True, antialias=True, background=(0.2, 0.1, 0.05, 0.5)),
    dict(linearise=True, filter3d=False, antialias=False, background=(0.0, 0.0, 0.0, 1.0)),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_forward_and_gradients_against_the_fp64_oracle(rast, case):
    c = CASES[case]
    p = scenes.diff_params(1024, 11 + case, degree=3, log_scale=(-3.0, 0.4))
    n = p["means"].shape[0]
    f3 = np.full(n, 0.008, np.float32) if c["filter3d"] else None
    s = raster.RenderSettings(width=W, height=HGT, linearise=c["linearise"], filter3d=c["filter3d"],
                              antialias=c["antialias"], background=c["background"])
    g = np.random.default_rng(100 + case).normal(0.0, 1.0, (HGT, W, 4))
    print(f"case {case}: {c}", flush=True)
    img, grads = _gpu_grads(rast, p, s, g, f3)
    frame = _oracle_frame(s, n, 15)
    ref_img = O.diff_render_f64(frame, p, filter3d=f3)
    d = np.abs(img.astype(np.float64) - ref_img).reshape(-1)
    print(f"  image vs FP64 brute force: p99 {np.quantile(d, 0.99):.2e} max {d.max():.2e}")
    assert np.quantile(d, 0.99) <= 1e-3 and d.max() <= 1e-2
    ref = D.gradients(frame, p, g, filter3d=f3)
    for name in GROUPS:
        a, b = grads[name].astype(np.float64).reshape(-1), ref[name].reshape(-1)
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        print(f"  {name:15s} relative error vs FP64 autodiff {rel:.2e}")
        assert rel < 1e-2, (name, rel)


def test_gradients_against_fp64_central_differences(rast):
    """The 8 largest gradient components of each group against FP64 central
    differences of the C brute force. The forward is piecewise smooth: a pixel
    whose power crosses a splat's cutoff (capped at 9, where alpha is still
    above 1/255), the alpha floor or the T < 1e-4 stop makes a step. A
    difference whose two ends change one of those choices (D.decisions) is
    excluded, as shown to straddle a step; the rest must agree to 1e-3 at worst."""
    p = scenes.diff_params(1024, 21, degree=3, log_scale=(-3.0, 0.4))
    n = p["means"].shape[0]
    s = raster.RenderSettings(width=W, height=HGT)
    g = np.random.default_rng(7).normal(0.0, 1.0, (HGT, W, 4))
    _, grads = _gpu_grads(rast, p, s, g)
    frame = _oracle_frame(s, n, 15)
    base_choices = D.decisions(frame, p)

    def loss(q):
        return float((O.diff_render_f64(frame, q) * g).sum())

    excluded = 0
    for name in GROUPS:
        flat = grads[name].reshape(-1)
        rel, steps = [], []
        for idx in np.argsort(-np.abs(flat))[:8]:
            base = p[name].reshape(-1)[idx]
            h = 1e-5 * max(1.0, abs(float(base)))
            plus, minus = {k: v.copy() for k, v in p.items()}, {k: v.copy() for k, v in p.items()}
            plus[name].reshape(-1)[idx] = np.float32(base + h)
            minus[name].reshape(-1)[idx] = np.float32(base - h)
            if not (D.same_decisions(D.decisions(frame, plus), base_choices) and
                    D.same_decisions(D.decisions(frame, minus), base_choices)):
                steps.append(int(idx))
                continue
            step = float(plus[name].reshape(-1)[idx]) - float(minus[name].reshape(-1)[idx])
            fd = (loss(plus) - loss(minus)) / step
            rel.append(abs(float(flat[idx]) - fd) / abs(fd))
        excluded += len(steps)
        print(f"  {name:15s} vs FP64 central differences: {len(rel)} smooth, max {np.max(rel):.2e} "
              f"median {np.median(rel):.2e}; straddling a step: {steps}")
        assert len(rel) >= 6, (name, steps)
        assert np.max(rel) < 1e-3, (name, rel)
    assert excluded <= 4


def test_autograd_wrapper_matches_the_c_abi(rast):
    p = scenes.diff_params(1024, 31, degree=3, log_scale=(-3.0, 0.4))
    s = raster.RenderSettings(width=W, height=HGT)
    g = np.random.default_rng(3).normal(0.0, 1.0, (HGT, W, 4))
    _, direct = _gpu_grads(rast, p, s, g)
    leaves = {k: v.clone().requires_grad_(True) for k, v in _device(p).items()}
    colour = rasterize(leaves["means"], leaves["log_scales"], leaves["quats"], leaves["opacity_logits"],
                       leaves["sh"], rast, _camera(), s)
    (colour * torch.from_numpy(g.astype(np.float32)).cuda()).sum().backward()
    for name in GROUPS:
        np.testing.assert_array_equal(leaves[name].grad.cpu().numpy(), direct[name])


def test_a_descent_step_lowers_the_l2_loss(rast):
    p = scenes.diff_params(1024, 41, degree=3, log_scale=(-3.0, 0.4))
    target_p = scenes.perturbed(p, seed=2)
    s = raster.RenderSettings(width=W, height=HGT)
    dev_t = _device(target_p)
    tgt = raster.RenderTargets(W, HGT)
    rast.render_diff(_camera(), raster.DiffSplats(dev_t["means"], dev_t["log_scales"], dev_t["quats"],
                                                  dev_t["opacity_logits"], dev_t["sh"]), s, tgt)
    target = tgt.colour.clone()
    leaves = {k: v.clone().requires_grad_(True) for k, v in _device(p).items()}

    def loss():
        out = rasterize(leaves["means"], leaves["log_scales"], leaves["quats"], leaves["opacity_logits"],
                        leaves["sh"], rast, _camera(), s)
        return ((out - target) ** 2).sum()

    before = loss()
    before.backward()
    with torch.no_grad():
        for k in GROUPS:
            leaves[k] -= 1e-4 * leaves[k].grad
    after = loss()
    assert after.item() < before.item()


def test_backward_needs_a_differentiable_frame(rast):
    p = scenes.diff_params(64, 5, degree=0)
    s = raster.RenderSettings(width=32, height=32)
    dev = _device(p)
    grads = {k: torch.empty_like(v) for k, v in dev.items()}
    cfg = scenes.workload("plain-fwd-65k-640x480", count=256)
    rast.render(_camera(), [raster.SplatInstance(cfg.scene.to_device())], s, raster.RenderTargets(32, 32))
    with pytest.raises(LrtError) as err:
        rast.backward(torch.zeros((32, 32, 4), device="cuda"), grads)
    assert err.value.code == 3
    # The backward walks 16-pixel tiles: other tile sizes are refused up front.
    splats = raster.DiffSplats(dev["means"], dev["log_scales"], dev["quats"], dev["opacity_logits"], dev["sh"])
    for tile in (8, 32):
        with pytest.raises(LrtError) as err:
            rast.render_diff(_camera(), splats, raster.RenderSettings(width=32, height=32, tile_size=tile),
                             raster.RenderTargets(32, 32))
        assert err.value.name == "Unsupported", err.value.name
    with pytest.raises(LrtError) as err:
        rast.render(_camera(), [raster.SplatInstance(cfg.scene.to_device())],
                    raster.RenderSettings(width=32, height=32, tile_size=12), raster.RenderTargets(32, 32))
    assert err.value.code == 3


def test_full_size_1m_720p_on_a_pixel_subset(rast):
    """The 1M-splat 720p fwd+bwd workload at its full size, on 64 pixels the
    cloud covers (coverage above 0.5, seeded): the HIP colours there against the
    FP64 brute force in the tile path's order (p99 <= 1e-3, max <= 1e-2), and
    the HIP gradients of sum over those pixels of colour * G against FP64
    central differences, for the four largest components of each group. The
    differences are taken over the splats that reach the 64 pixels (their
    colours there are the whole cloud's: the walk keeps their order); a
    component whose difference moves by more than 1e-2 between the steps h and
    h/4 straddles a step of the forward and is left out (at most one a group),
    the rest must agree to 1e-3 (measured: 3e-6 at worst; the image 2e-6)."""
    n, width, height = 1_000_000, 1280, 720
    p = scenes.diff_params(n, 0, degree=3)
    s = raster.RenderSettings(width=width, height=height)
    dev = _device(p)
    splats = raster.DiffSplats(dev["means"], dev["log_scales"], dev["quats"], dev["opacity_logits"], dev["sh"])
    targets = raster.RenderTargets(width, height)
    rast.render_diff(_camera(), splats, s, targets)
    torch.cuda.synchronize()
    img = targets.colour.cpu().numpy()
    rng = np.random.default_rng(5)
    ys, xs = np.nonzero(img[..., 3] > 0.5)
    pick = rng.choice(xs.size, 64, replace=False)
    px, py = xs[pick].astype(np.uint32), ys[pick].astype(np.uint32)
    frame = _oracle_frame(s, n, 15)
    ref = O.diff_render_f64_pixels(frame, p, px, py, keyed=True)
    d = np.abs(img[py, px].astype(np.float64) - ref).reshape(-1)
    print(f"1M splats, 64 pixels vs FP64 brute force: p99 {np.quantile(d, 0.99):.2e} max {d.max():.2e}", flush=True)
    assert np.quantile(d, 0.99) <= 1e-3 and d.max() <= 1e-2

    g = np.zeros((height, width, 4))
    g[py, px] = rng.normal(0.0, 1.0, (64, 4))
    grads = {k: torch.empty_like(v) for k, v in dev.items()}
    rast.backward(torch.from_numpy(g.astype(np.float32)).cuda(), grads)
    torch.cuda.synchronize()
    grads = {k: v.cpu().numpy() for k, v in grads.items()}

    # The splats that reach any of the 64 pixels (their power there within the
    # cutoff plus a margin the differences' steps cannot cross), in index order.
    t = {k: torch.from_numpy(np.ascontiguousarray(p[key], dtype=np.float32))
         for k, key in (("means", "means"), ("log_scales", "log_scales"), ("quats", "quats"),
                        ("logits", "opacity_logits"), ("sh", "sh"))}
    rec = D.project(frame, t["means"], t["log_scales"], t["quats"], t["logits"], t["sh"])
    reach = np.zeros(n, bool)
    for x, y in zip(px, py):
        dx = float(x) + 0.5 - rec["cx"]
        dy = float(y) + 0.5 - rec["cy"]
        power = rec["a"] * dx * dx + 2.0 * rec["b"] * dx * dy + rec["c"] * dy * dy
        reach |= (rec["valid"] & (power <= rec["cutoff"] + 1.0)).numpy()
    keep = np.nonzero(reach)[0]
    sub = {k: np.ascontiguousarray(v[keep]) for k, v in p.items()}
    sub_frame = _oracle_frame(s, keep.size, 15)
    gsub = g[py, px]
    assert np.abs(O.diff_render_f64_pixels(sub_frame, sub, px, py) - ref).max() < 1e-12
    print(f"  {keep.size} splats reach the 64 pixels", flush=True)

    def loss(q):
        return float((O.diff_render_f64_pixels(sub_frame, q, px, py) * gsub).sum())

    position = {int(k): j for j, k in enumerate(keep)}
    for name in GROUPS:
        flat = grads[name].reshape(-1)
        per = flat.size // n
        rel, steps = [], []
        for idx in np.argsort(-np.abs(flat))[:4]:
            j = position[int(idx) // per]
            at = j * per + int(idx) % per
            base = float(sub[name].reshape(-1)[at])
            fd = []
            for h in (1e-4 * max(1.0, abs(base)), 2.5e-5 * max(1.0, abs(base))):
                plus, minus = {k: v.copy() for k, v in sub.items()}, {k: v.copy() for k, v in sub.items()}
                plus[name].reshape(-1)[at] = np.float32(base + h)
                minus[name].reshape(-1)[at] = np.float32(base - h)
                step = float(plus[name].reshape(-1)[at]) - float(minus[name].reshape(-1)[at])
                fd.append((loss(plus) - loss(minus)) / step)
            if abs(fd[0] - fd[1]) > 1e-2 * abs(fd[1]):
                steps.append(int(idx))
                continue
            rel.append(abs(float(flat[idx]) - fd[1]) / abs(fd[1]))
        print(f"  {name:15s} vs FP64 central differences: {len(rel)} smooth, max {np.max(rel):.2e}; "
              f"straddling a step: {steps}", flush=True)
        assert len(rel) >= 3, (name, steps)
        assert np.max(rel) < 1e-3, (name, rel)
