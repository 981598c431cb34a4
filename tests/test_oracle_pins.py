"""Pins two parts of the oracle that nothing else fixes independently, and shows
that each pin fails on a plausible mistake planted in the oracle.

1. The Mip-Splatting 3D filter (Yu et al., "Mip-Splatting", CVPR 2024, eq. 9):
   the 3D covariance becomes Sigma + f^2 I and the opacity is scaled by
   sqrt(det Sigma / det(Sigma + f^2 I)). The check rebuilds the 2D covariance
   from the packed splat in float64 with full 3x3 matrices (scipy's rotation of
   the decoded quaternion, numpy determinants and inverses): for an isotropic
   splat under an orthographic camera it is the closed form
   focal^2 (s^2 + f^2) I + dilation I with opacity factor (s^2/(s^2+f^2))^(3/2);
   for a rotated anisotropic splat under a perspective camera it is
   J W (R diag(s^2) R^T + f^2 I) W^T J^T + dilation I.
2. The exact tile count (frame.slang:113-138): a splat touches a tile when its
   cutoff ellipse {a x^2 + 2 b x y + c y^2 <= cutoff} reaches the box spanned
   by the tile's pixel centres. The check decides that in float64 by another
   route than the oracle's clamped minimiser: the ellipse meets the box when
   the centre lies inside or one of the box's four edges crosses it (the roots
   of the quadratic along the edge overlap the edge). Every tile the emitted
   pairs hold must be reached at cutoff (1 + eps), and every tile reached at
   cutoff (1 - eps) must be held: an over-count and an under-count both fail.
"""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation

import fixtures as F
import harness as H
import mutants
import oracle_lib as O
from lucabrtrender_amd import raster

# ---------------------------------------------------------------------------
# 1. the 3D filter
# ---------------------------------------------------------------------------


def _project_one(lib, frame, scene):
    """One splat through `lib`'s or_project: (record [12] float32, touched)."""
    proj = np.zeros((1, 12), np.float32)
    rects = np.zeros((1, 4), np.uint32)
    touched = np.zeros(1, np.uint32)
    visible = np.zeros(1, np.uint32)
    keys = np.zeros(1, np.uint32)
    lib.or_project(frame, O.ptr(scene.positions), O.ptr(scene.shape), O.ptr(scene.sh), O.ptr(scene.filter3d),
                   O.ptr(proj), O.ptr(rects), O.ptr(touched), O.ptr(visible), O.ptr(keys), 1)
    return proj[0], int(touched[0])


def _splat(position, opacity, scales, wxyz, filter_size):
    b = F.CloudBuilder()
    b.add(*position, opacity, *scales, wxyz, (0.6, 0.4, 0.2))
    scene = b.decode(0)
    scene.filter3d = np.array([filter_size], np.float32)
    return scene


def _decoded(scene):
    """The packed splat's own values in float64: scales exp(f16 log s), the
    decoded quaternion's rotation, the linear opacity."""
    halves = scene.shape[0, 1:3].view(np.float16).astype(np.float64)      # ln sx, ln sy, ln sz, base r
    s = np.exp(halves[:3])
    q = (O.C.c_float * 4)()
    O.lib.or_decode_quaternion(int(scene.shape[0, 0]), q)
    rot = Rotation.from_quat([q[0], q[1], q[2], q[3]]).as_matrix()          # (x, y, z, w), scalar last
    return s, rot, float(scene.positions[0, 3])


def _expected(frame, scene, settings):
    """(2D covariance [2, 2], alpha) from the splat in float64."""
    s, rot, opacity = _decoded(scene)
    f = float(scene.filter3d[0])
    sigma = rot @ np.diag(s * s) @ rot.T
    filtered = sigma + f * f * np.eye(3)
    opacity = opacity * np.sqrt(np.linalg.det(sigma) / np.linalg.det(filtered))
    m = np.array(frame.m[:], np.float64).reshape(3, 4)
    t = m[:, :3] @ scene.positions[0, :3].astype(np.float64) + m[:, 3]
    if frame.orthographic:
        jac = np.array([[frame.focal_x, 0.0, 0.0], [0.0, frame.focal_y, 0.0]])
    else:
        jac = np.array([[frame.focal_x / t[2], 0.0, -frame.focal_x * t[0] / t[2] ** 2],
                        [0.0, frame.focal_y / t[2], -frame.focal_y * t[1] / t[2] ** 2]])
    a = jac @ m[:, :3]
    cov = a @ filtered @ a.T
    dilated = cov + settings.dilation * np.eye(2)
    alpha = opacity * (np.sqrt(np.linalg.det(cov) / np.linalg.det(dilated)) if settings.antialias else 1.0)
    return dilated, alpha


CASES = {
    # isotropic, orthographic: the closed form of the module docstring
    "isotropic-ortho": dict(position=(0.05, -0.03, 0.0), scales=(0.1, 0.1, 0.1), wxyz=(1, 0, 0, 0), f=0.08,
                            eye=(0.0, 0.0, 5.0), lens=dict(orthographic=True, focal=20.0), dilation=0.3),
    # rotated anisotropic, perspective, the headline scene's 2D filter 0.1
    "rotated-persp": dict(position=(0.1, 0.05, -0.2), scales=(0.15, 0.04, 0.012), wxyz=(0.8, 0.3, -0.4, 0.35),
                          f=0.03, eye=(0.6, 0.4, 4.0), lens=dict(focal=35.0), dilation=0.1),
    "rotated-ortho-noaa": dict(position=(0.0, 0.0, 0.0), scales=(0.02, 0.2, 0.07), wxyz=(0.2, -0.7, 0.1, 0.6),
                               f=0.05, eye=(1.0, 2.0, 4.0), lens=dict(orthographic=True, focal=15.0),
                               dilation=0.3, antialias=False),
}


def _filter_case(lib, name):
    c = CASES[name]
    scene = _splat(c["position"], 0.7, c["scales"], c["wxyz"], c["f"])
    cam = raster.Camera.looking_at(c["eye"], (0.0, 0.0, 0.0))
    cam.lens = raster.Lens(**c["lens"])
    s = raster.RenderSettings(width=160, height=120, dilation=c["dilation"], filter3d=True,
                              antialias=c.get("antialias", True), max_sh_degree=0)
    frames, _, _ = H.oracle_frames(cam, s, [(scene, np.eye(4), raster.SplatEdit())])
    rec, touched = _project_one(lib, frames[0], scene)
    assert touched > 0, name
    cov_expected, alpha_expected = _expected(frames[0], scene, s)
    conic = np.array([[rec[4], rec[5]], [rec[5], rec[6]]], np.float64)
    cov = np.linalg.inv(conic)
    cov_err = np.abs(cov - cov_expected).max() / np.abs(cov_expected).max()
    alpha_err = abs(float(rec[7]) - alpha_expected) / alpha_expected
    return cov_err, alpha_err, cov_expected, alpha_expected


@pytest.mark.parametrize("name", list(CASES))
def test_3d_filter_matches_the_float64_covariance_and_opacity(name):
    cov_err, alpha_err, cov, alpha = _filter_case(O.lib, name)
    print(f"{name}: cov {np.round(cov, 3).tolist()} rel err {cov_err:.1e}; alpha {alpha:.5f} rel err {alpha_err:.1e}")
    assert cov_err < 2e-5 and alpha_err < 2e-5


def test_3d_filter_isotropic_closed_form():
    """The isotropic orthographic case written out: focal^2 (s^2 + f^2) + dilation
    on the diagonal, no off-diagonal, and alpha = o (s^2/(s^2+f^2))^(3/2) times
    the antialiasing factor det(cov)/det(cov + dilation) to the half."""
    c = CASES["isotropic-ortho"]
    cov_err, alpha_err, cov, alpha = _filter_case(O.lib, "isotropic-ortho")
    scene = _splat(c["position"], 0.7, c["scales"], c["wxyz"], c["f"])
    s, _, o = _decoded(scene)
    assert np.allclose(s, s[0], rtol=1e-12)        # all three f16 logs the same
    s2 = s[0] ** 2
    cam = raster.Camera.looking_at(c["eye"], (0.0, 0.0, 0.0))
    cam.lens = raster.Lens(**c["lens"])
    focal = raster.projection_for(cam, 160, 120).focal_x
    v = np.float32(focal) ** 2 * (s2 + np.float32(c["f"]) ** 2)
    closed_cov = np.diag([v + c["dilation"], v + c["dilation"]])
    closed_alpha = o * (s2 / (s2 + np.float32(c["f"]) ** 2)) ** 1.5 * (v / (v + c["dilation"]))
    # the float64 matrix route agrees with the closed form (its camera rows are
    # the float32 frame's, hence 1e-7) ...
    assert np.allclose(cov, closed_cov, rtol=1e-7, atol=1e-9)
    assert abs(alpha - closed_alpha) / closed_alpha < 1e-7
    # ... and the oracle's record agrees with both
    frames, _, _ = H.oracle_frames(cam, raster.RenderSettings(width=160, height=120, dilation=c["dilation"],
                                                              filter3d=True, max_sh_degree=0),
                                   [(scene, np.eye(4), raster.SplatEdit())])
    rec, _ = _project_one(O.lib, frames[0], scene)
    ora_cov = np.linalg.inv(np.array([[rec[4], rec[5]], [rec[5], rec[6]]], np.float64))
    assert np.allclose(ora_cov, closed_cov, rtol=2e-5, atol=1e-3)
    assert abs(float(rec[7]) - closed_alpha) / closed_alpha < 2e-5
    assert cov_err < 2e-5 and alpha_err < 2e-5


FILTER_MISTAKES = {
    "filter-not-squared": [("const float f2 = f * f;", "const float f2 = f;")],
    "filter-ratio-inverted": [("*opacity = *opacity * sqrtf(det1 / det2);", "*opacity = *opacity * sqrtf(det2 / det1);")],
    "filter-one-axis-dropped": [("*sz = sqrtf(a2z);", "*sz = sqrtf(s2z);")],
    "filter-no-opacity": [("*opacity = *opacity * sqrtf(det1 / det2);", "*opacity = *opacity;")],
}


@pytest.fixture(scope="module")
def mutant_dir(tmp_path_factory):
    return tmp_path_factory.mktemp("mutants")


@pytest.mark.parametrize("mistake", list(FILTER_MISTAKES))
def test_3d_filter_pin_catches_a_planted_mistake(mutant_dir, mistake):
    lib = mutants.build(mutant_dir, mistake, FILTER_MISTAKES[mistake])
    worst = 0.0
    for name in CASES:
        cov_err, alpha_err, _, _ = _filter_case(lib, name)
        worst = max(worst, cov_err, alpha_err)
    print(f"{mistake}: worst relative error {worst:.2e}")
    assert worst > 1e-3


# ---------------------------------------------------------------------------
# 2. the exact tile count
# ---------------------------------------------------------------------------


def _tile_boxes(frame):
    """Pixel-centre boxes of every tile, in tile order (ty * tiles_x + tx)."""
    tx = np.arange(frame.tiles_x, dtype=np.float64)
    ty = np.arange(frame.tiles_y, dtype=np.float64)
    x0, y0 = tx * 16 + 0.5, ty * 16 + 0.5
    x1 = np.minimum((tx + 1) * 16, frame.width) - 1 + 0.5
    y1 = np.minimum((ty + 1) * 16, frame.height) - 1 + 0.5
    gx0, gy0 = np.meshgrid(x0, y0)
    gx1, gy1 = np.meshgrid(x1, y1)
    return gx0.reshape(-1), gx1.reshape(-1), gy0.reshape(-1), gy1.reshape(-1)


def ellipse_meets_boxes(a, b, c, cx, cy, thr, boxes):
    """[splats, tiles] bool: does {a x^2 + 2 b x y + c y^2 <= thr} about (cx, cy)
    meet each box? float64; a, c > 0 and a c > b^2 (a conic of a splat)."""
    a, b, c, cx, cy, thr = (np.asarray(v, np.float64)[:, None] for v in (a, b, c, cx, cy, thr))
    x0, x1, y0, y1 = (v[None, :] for v in boxes)
    x0, x1, y0, y1 = x0 - cx, x1 - cx, y0 - cy, y1 - cy
    inside = (x0 <= 0) & (x1 >= 0) & (y0 <= 0) & (y1 >= 0)
    det = a * c - b * b

    def along_x(x):          # the edge at this x, y in [y0, y1]: c y^2 + 2 b x y + a x^2 - thr <= 0
        disc = c * thr - x * x * det
        r = np.sqrt(np.maximum(disc, 0.0))
        return (disc >= 0) & ((-b * x - r) / c <= y1) & ((-b * x + r) / c >= y0)

    def along_y(y):          # the edge at this y, x in [x0, x1]
        disc = a * thr - y * y * det
        r = np.sqrt(np.maximum(disc, 0.0))
        return (disc >= 0) & ((-b * y - r) / a <= x1) & ((-b * y + r) / a >= x0)

    return inside | along_x(x0) | along_x(x1) | along_y(y0) | along_y(y1)


EPS = 1e-5


def count_check(r):
    """(tiles held but not reached even at cutoff (1 + EPS), tiles reached at
    cutoff (1 - EPS) but not held, ambiguous tiles, held tiles, anisotropic
    drawn splats) for one oracle frame `r` of one instance."""
    f = r["frames"][0]
    drawn = np.nonzero(r["touched"])[0]
    proj = r["proj"][drawn].astype(np.float64)
    held = np.zeros((drawn.size, f.tiles_x * f.tiles_y), bool)
    row = np.full(r["touched"].size, -1, np.int64)
    row[drawn] = np.arange(drawn.size)
    held[row[r["pair_splats"]], r["pair_tiles"]] = True
    assert np.array_equal(held.sum(axis=1), r["touched"][drawn])
    boxes = _tile_boxes(f)
    a, b, c, cx, cy, cut = proj[:, 4], proj[:, 5], proj[:, 6], proj[:, 0], proj[:, 1], proj[:, 3]
    upper = ellipse_meets_boxes(a, b, c, cx, cy, cut * (1 + EPS), boxes)
    lower = ellipse_meets_boxes(a, b, c, cx, cy, cut * (1 - EPS), boxes)
    over = int((held & ~upper).sum())
    under = int((lower & ~held).sum())
    aniso = int((np.abs(b) > 0.05 * np.sqrt(a * c)).sum())
    return over, under, int((upper & ~lower).sum()), int(held.sum()), aniso


def _count_frames():
    cloud = F.random_cloud(3000, 7).decode(3)
    big = F.random_cloud(300, 11, min_size=0.05, max_size=1.2).decode(0)
    persp = raster.Camera.looking_at((1.5, 1.0, 6.0), (0, 0, 0))
    persp.lens = raster.Lens(focal=30.0)
    ortho = raster.Camera.looking_at((0.0, 3.0, 5.0), (0, 0, 0))
    ortho.lens = raster.Lens(orthographic=True, focal=5.0)
    edge = raster.Camera.looking_at((0.3, -0.2, 2.2), (0, 0, 0))
    edge.lens = raster.Lens(focal=18.0)
    return [
        ("random-3000-persp", cloud, persp, raster.RenderSettings(width=250, height=190)),
        ("random-3000-ortho-dil0.1", cloud, ortho, raster.RenderSettings(width=250, height=190, dilation=0.1)),
        ("large-300-frustum-edge", big, edge, raster.RenderSettings(width=203, height=157, max_sh_degree=0)),
    ]


def _count_case(lib, name, scene, cam, s):
    frames, clouds, n = H.oracle_frames(cam, s, [(scene, np.eye(4), raster.SplatEdit())])
    r = O.raster(frames, clouds, n, want_image=False, using=lib)
    r["frames"] = frames
    return count_check(r)


@pytest.mark.parametrize("case", range(3))
def test_tile_count_is_two_sided_exact(case):
    name, scene, cam, s = _count_frames()[case]
    over, under, ambiguous, held, aniso = _count_case(O.lib, name, scene, cam, s)
    print(f"{name}: {held} tiles held, {aniso} anisotropic splats, {ambiguous} within {EPS} of the cutoff")
    assert aniso > 50
    assert over == 0 and under == 0
    assert ambiguous <= held * 1e-3


COUNT_MISTAKES = {
    "count-whole-rectangle": [("return min_power_over_box(a, b, c, fx - cx, fy - cy, lx - cx, ly - cy) <= cutoff;",
                               "return 1;")],
    "count-clamp-sign": [("const float y = c > 1e-12f ? clampf_(-b * x / c, loy, hiy) : loy;",
                          "const float y = c > 1e-12f ? clampf_(b * x / c, loy, hiy) : loy;")],
    "count-box-one-pixel-wide": [("const float lx = (float)(ex - 1u) + 0.5f;", "const float lx = (float)ex + 0.5f;")],
    "count-cross-term-sign": [("best = fmin_(best, a * x * x + 2.0f * b * x * y + c * y * y);\n    }\n    for",
                               "best = fmin_(best, a * x * x - 2.0f * b * x * y + c * y * y);\n    }\n    for")],
}


@pytest.mark.parametrize("mistake", list(COUNT_MISTAKES))
def test_tile_count_pin_catches_a_planted_mistake(mutant_dir, mistake):
    lib = mutants.build(mutant_dir, mistake, COUNT_MISTAKES[mistake])
    wrong = 0
    for name, scene, cam, s in _count_frames():
        over, under, _, held, _ = _count_case(lib, name, scene, cam, s)
        print(f"{mistake} on {name}: {over} over, {under} under of {held}")
        wrong += over + under
    assert wrong > 0
