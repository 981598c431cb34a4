"""The FP64 autodiff oracle of the backward (oracle/diff_oracle.py), pinned on
the CPU before it judges the HIP gradients:

- its forward agrees with the C restatement of the brute force
  (or_diff_render_f64, lrt_oracle.c) to FP64 rounding, so both restate the
  same function;
- its autodiff gradients agree with central finite differences of that C
  forward in FP64, parameter by parameter, for every parameter group;
- a planted mistake (a transposed quaternion in the C forward's input) is
  caught by the first check.

The reference has no backward to compare with; these checks are what pins
the gradient oracle.
"""
import numpy as np
import pytest
import torch

import oracle_lib as O
from lucabrtrender_amd import scenes
from oracle import diff_oracle as D

W, HGT = 64, 48
GROUPS = ("means", "log_scales", "quats", "opacity_logits", "sh")


def _frame(count, rest, linearise=True, filter3d=False, antialias=True, background=(0, 0, 0, 0)):
    cam = O.camera(O.look_at((0.0, 0.0, 0.0), (0.0, 0.0, -1.0)), focal=scenes.focal_for_fov(60.0))
    proj = O.projection_for(cam, W, HGT)
    s = O.settings(W, HGT, antialias=antialias, linearise=linearise, filter3d=filter3d, background=background)
    return O.diff_frame(proj, s, count, rest)


def _scene(count=300, seed=3):
    # Larger than the plain cube's splats so that a 64x48 frame resolves their shapes.
    return scenes.diff_params(count, seed, degree=3, log_scale=(-2.6, 0.4))


def _torch_params(p):
    return [torch.from_numpy(np.asarray(p[k], dtype=np.float64)) for k in GROUPS]


@pytest.mark.parametrize("linearise,filter3d,antialias", [(True, False, True), (False, True, True),
                                                          (True, False, False)])
def test_torch_forward_matches_the_c_restatement(linearise, filter3d, antialias):
    p = _scene()
    f3 = np.full(p["means"].shape[0], 0.01, np.float32) if filter3d else None
    frame = _frame(p["means"].shape[0], 15, linearise=linearise, filter3d=filter3d, antialias=antialias,
                   background=(0.1, 0.2, 0.3, 0.5))
    c_img = O.diff_render_f64(frame, p, filter3d=f3)
    t_img = D.render(frame, *_torch_params(p), filter3d=torch.from_numpy(f3.astype(np.float64)) if filter3d else None)
    assert np.abs(c_img).max() > 0.05            # the frame draws something
    np.testing.assert_allclose(t_img.numpy(), c_img, rtol=0, atol=1e-10)


def test_a_planted_quaternion_transpose_is_caught():
    p = _scene()
    frame = _frame(p["means"].shape[0], 15)
    bad = dict(p)
    bad["quats"] = p["quats"][:, [3, 0, 1, 2]].copy()   # w first: the wrong convention
    c_img = O.diff_render_f64(frame, bad)
    t_img = D.render(frame, *_torch_params(p)).numpy()
    assert np.abs(t_img - c_img).max() > 1e-3


def _loss(frame, p, g, f3=None):
    return float((O.diff_render_f64(frame, p, filter3d=f3) * g).sum())


@pytest.mark.parametrize("linearise,filter3d", [(True, False), (False, True)])
def test_autodiff_matches_central_differences(linearise, filter3d):
    p = _scene()
    n = p["means"].shape[0]
    f3 = np.full(n, 0.01, np.float32) if filter3d else None
    frame = _frame(n, 15, linearise=linearise, filter3d=filter3d)
    g = np.random.default_rng(5).normal(0.0, 1.0, (HGT, W, 4))
    grads = D.gradients(frame, p, g, filter3d=f3)
    checked = 0
    for name in GROUPS:
        ad = grads[name].reshape(-1)
        # the components with the largest gradients, where a relative error means something
        picks = np.argsort(-np.abs(ad))[:12]
        rel = []
        for idx in picks:
            base = p[name].reshape(-1)[idx]
            h = 1e-5 * max(1.0, abs(float(base)))
            plus, minus = {k: v.copy() for k, v in p.items()}, {k: v.copy() for k, v in p.items()}
            plus[name].reshape(-1)[idx] = np.float32(base + h)
            minus[name].reshape(-1)[idx] = np.float32(base - h)
            step = float(plus[name].reshape(-1)[idx]) - float(minus[name].reshape(-1)[idx])
            fd = (_loss(frame, plus, g, f3) - _loss(frame, minus, g, f3)) / step
            rel.append(abs(fd - ad[idx]) / max(abs(ad[idx]), 1e-12))
            checked += 1
        rel = np.array(rel)
        # Steps of 1e-5 relative cross no pixel's skip threshold on this scene;
        # the agreement is at the level of the differences' own rounding.
        assert np.median(rel) < 1e-7 and rel.max() < 1e-5, (name, rel)
    assert checked == 12 * len(GROUPS)


def test_the_step_detector_flags_the_difference_that_straddles_a_cutoff():
    """The differentiable frame's FD sub-scene at 128 x 96: the y of splat 18's
    mean sits 4.3e-5 in power outside the 3-sigma cutoff cap (power 9) at
    pixel (64, 31), where its alpha is still 2.3/255. A step of -1e-5 brings
    that pixel in: D.decisions must change there and only there, the one-sided
    difference on that side must grow as 1/h (a jump, not a slope), and on the
    smooth side it must equal the autodiff gradient."""
    w, h = 128, 96
    p = scenes.diff_params(1024, 21, degree=3, log_scale=(-3.0, 0.4))
    cam = O.camera(O.look_at((0.0, 0.0, 0.0), (0.0, 0.0, -1.0)), focal=scenes.focal_for_fov(60.0))
    frame = O.diff_frame(O.projection_for(cam, w, h), O.settings(w, h), 1024, 15)
    g = np.random.default_rng(7).normal(0.0, 1.0, (h, w, 4))
    base_choices = D.decisions(frame, p)
    ad = float(D.gradients(frame, p, g)["means"].reshape(-1)[55])

    def moved(step):
        q = {k: v.copy() for k, v in p.items()}
        q["means"].reshape(-1)[55] = np.float32(p["means"].reshape(-1)[55] + step)
        return q

    def loss(q):
        return float((O.diff_render_f64(frame, q) * g).sum())

    down = D.decisions(frame, moved(-1e-5))
    assert np.array_equal(down[0], base_choices[0])
    flips = np.argwhere(down[1] != base_choices[1])
    assert [(int(a) % w, int(a) // w, int(base_choices[0][b])) for a, b in flips] == [(64, 31, 18)]
    assert D.same_decisions(D.decisions(frame, moved(1e-5)), base_choices)
    l0 = loss(p)
    jumps = []
    for step in (1e-5, 1e-6):
        lo, hi = moved(-step), moved(step)
        below = (l0 - loss(lo)) / (float(p["means"].reshape(-1)[55]) - float(lo["means"].reshape(-1)[55]))
        above = (loss(hi) - l0) / (float(hi["means"].reshape(-1)[55]) - float(p["means"].reshape(-1)[55]))
        assert abs(above - ad) < 1e-3 * abs(ad), (above, ad)
        jumps.append((below - ad) * step)
    # the same jump in the loss at both step sizes: a step of the forward
    assert abs(jumps[0]) > 1e-3 and abs(jumps[1] / jumps[0] - 1.0) < 0.05, jumps
