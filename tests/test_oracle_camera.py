"""Pins the oracle's camera and projection: the reference's closed-form single
splat (tests/render/test_splat_render.cpp:123-150), the EWA covariance against
an independent double-precision matrix product (exact for an orthographic
camera), and the library's host camera math bit-identical to the oracle's."""
import numpy as np
import pytest

import fixtures as F
import harness as H
import oracle_lib as O
from lucabrtrender_amd import raster


def _project(scene, cam, s):
    frames, clouds, n = H.oracle_frames(cam, s, [(scene, np.eye(4), raster.SplatEdit())])
    total = scene.count
    proj = np.zeros((total, 12), np.float32)
    touched = np.zeros(total, np.uint32)
    O.lib.or_project(frames, O.ptr(scene.positions), O.ptr(scene.shape), O.ptr(scene.sh) if scene.sh is not None
                     else None, None, O.ptr(proj), None, O.ptr(touched), None, None, 1)
    return proj, touched, frames[0]


def _quat(word):
    import ctypes as C
    q = (C.c_float * 4)()
    O.lib.or_decode_quaternion(int(word), q)
    return np.array(q[:], np.float64)


def _rotation(q):
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def _stored_covariance(scene, i):
    s = scene.shape[i]
    ls = [O.lib.or_half_to_float(int(s[1]) & 0xFFFF), O.lib.or_half_to_float(int(s[1]) >> 16),
          O.lib.or_half_to_float(int(s[2]) & 0xFFFF)]
    scale = np.exp(np.array(ls, np.float64))
    r = _rotation(_quat(s[0]))
    m = r @ np.diag(scale)
    return m @ m.T


def test_single_splat_lands_where_the_camera_arithmetic_says():
    b = F.CloudBuilder()
    b.add(1.0, 0.5, -10.0, 0.9, 0.02, 0.02, 0.02, (1, 0, 0, 0), (1, 1, 1))
    scene = b.decode(3)
    cam = raster.Camera()
    cam.lens = raster.Lens(focal=24.576)             # fx = width
    s = raster.RenderSettings(width=320, height=240)
    out = O.raster(*H.oracle_frames(cam, s, [(scene, np.eye(4), raster.SplatEdit())]))
    alpha = out["colour"][..., 3]
    y, x = np.unravel_index(np.argmax(alpha), alpha.shape)
    # x: 160 + 320 * 1 / 10 = 192; y: 120 + 320 * 0.5 / 10 = 136, bottom row first
    assert abs(int(x) - 192) <= 1 and abs(int(y) - 136) <= 1


@pytest.mark.parametrize("angle", [0.0, 30.0, 77.0])
def test_orthographic_ewa_is_the_exact_covariance(angle):
    b = F.CloudBuilder()
    half = np.radians(angle) / 2
    b.add(0.3, -0.2, -5.0, 0.8, 0.3, 0.1, 0.05, (np.cos(half), 0.0, 0.0, np.sin(half)), (0.5, 0.5, 0.5))
    scene = b.decode(0)
    cam = raster.Camera()
    cam.lens = raster.Lens(orthographic=True, focal=5.0)
    s = raster.RenderSettings(width=200, height=160, antialias=False)
    proj, touched, f = _project(scene, cam, s)
    assert touched[0] > 0
    fx = 200 * 5.0 / 24.576
    cov = _stored_covariance(scene, 0)
    w = np.diag([1.0, 1.0, -1.0])                   # the view's z flip
    j = np.array([[fx, 0, 0], [0, fx, 0]])
    cov2 = j @ w @ cov @ w.T @ j.T + 0.3 * np.eye(2)
    conic = np.linalg.inv(cov2)
    got = proj[0, 4:7].astype(np.float64)
    assert np.allclose(got, [conic[0, 0], conic[0, 1], conic[1, 1]], rtol=2e-5, atol=1e-9)
    assert np.allclose(proj[0, :2], [fx * 0.3 + 100.0, fx * -0.2 + 80.0], atol=1e-4)


def test_perspective_ewa_against_the_jacobian_in_double():
    b = F.random_cloud(200, 11)
    scene = b.decode(3)
    cam = raster.Camera.looking_at((1.5, 1.0, 6.0), (0.0, 0.0, 0.0))
    cam.lens = raster.Lens(focal=30.0)
    s = raster.RenderSettings(width=250, height=190)
    proj, touched, f = _project(scene, cam, s)
    p = raster.projection_for(cam, 250, 190)
    wv = np.array(p.world_to_view[:]).reshape(4, 4)
    checked = 0
    for i in np.nonzero(touched)[0]:
        v = wv[:3, :3] @ scene.positions[i, :3].astype(np.float64) + wv[:3, 3]
        x, y, z = v
        if abs(x / z * p.focal_x) > 0.5 * 250 or abs(y / z * p.focal_y) > 0.5 * 190:
            continue                                  # the Jacobian clamp applies outside 1.3x the frustum
        jac = np.array([[p.focal_x / z, 0, -p.focal_x * x / z ** 2], [0, p.focal_y / z, -p.focal_y * y / z ** 2]])
        cov2 = jac @ wv[:3, :3] @ _stored_covariance(scene, i) @ wv[:3, :3].T @ jac.T
        det0 = np.linalg.det(cov2)
        cov2 += 0.3 * np.eye(2)
        conic = np.linalg.inv(cov2)
        alpha = scene.positions[i, 3] * np.sqrt(max(det0, 0) / np.linalg.det(cov2))
        assert np.allclose(proj[i, :2], [p.focal_x * x / z + p.centre_x, p.focal_y * y / z + p.centre_y],
                           atol=2e-3)
        assert np.allclose(proj[i, 4:7], [conic[0, 0], conic[0, 1], conic[1, 1]], rtol=2e-3, atol=1e-6)
        assert abs(proj[i, 7] - alpha) <= 2e-4 * max(alpha, 1e-3)
        assert proj[i, 3] == pytest.approx(min(2 * np.log(max(255 * proj[i, 7], 1.0001)), 9.0), rel=1e-5)
        checked += 1
    assert checked > 50


CAMERAS = [
    dict(eye=(1.5, 1.0, 6.0), target=(0, 0, 0), lens=dict(focal=30.0)),
    dict(eye=(0.3, -0.2, 2.2), target=(0, 0, 0), lens=dict(focal=18.0)),
    dict(eye=(0.0, 3.0, 5.0), target=(0, 0, 0), lens=dict(orthographic=True, focal=5.0)),
    dict(eye=(-2.0, 0.5, 5.0), target=(0, 0, 0),
         lens=dict(window_translate=(0.2, 0.0), window_scale=(1.0, 1.3), window_roll=25.0)),
    dict(eye=(2.0, 1.5, 7.0), target=(0, 0, -0.5), lens={}),
]


@pytest.mark.parametrize("c", CAMERAS)
def test_library_host_math_is_the_oracle_bit_for_bit(c):
    cam = raster.Camera.looking_at(c["eye"], c["target"])
    cam.lens = raster.Lens(**c["lens"])
    assert np.array_equal(np.array(cam.camera_to_world), O.look_at(c["eye"], c["target"]))
    p = raster.projection_for(cam, 250, 190)
    op = O.projection_for(H.oracle_camera(cam), 250, 190)
    assert bytes(p) == bytes(op)
    o2w = np.eye(4)
    o2w[:3, :3] = np.array([[0.77, 0, 0.64], [0, 0.7, 0], [-0.64, 0, 1.0]])
    o2w[:3, 3] = [0.5, 0.0, -1.5]
    s = raster.RenderSettings(width=250, height=190, antialias=False, dilation=0.1)
    mine = raster.frame_params_bytes(p, raster.SplatInstance(None, tuple(o2w.reshape(-1))), s, base=0, count=10,
                                     rest_per_colour=3)
    scene = type("S", (), dict(count=10, rest_per_colour=3, sh_words=5, positions=None, shape=None, sh=None,
                               filter3d=None))()
    frames, _, _ = O.frames_for(op, H.oracle_settings(s), [(scene, o2w, H.oracle_edit(raster.SplatEdit()))])
    assert bytes(frames[0]) == mine
