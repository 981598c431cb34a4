"""Pins the oracle's levels of detail (modules/lod/src/Lod.cpp:79-546,
shaders/lrt/lod/*.slang) with what the method fixes, not by retyping it:

- Morton codes against a bit-by-bit interleave written here;
- each level's groups against numpy's unique runs of the codes' prefixes;
- leaf moments against float64 sums of the decoded splats; merged moments
  against leaf moments taken directly at the coarser level (moments add);
- each merged Gaussian's covariance, mean, opacity and colour against the
  moments it came from (float64 eigen-decomposition), and a cell holding one
  splat merging back into that splat (tests/lod/test_lod.cpp:107-153);
- the cut drawing every place exactly once, at every threshold and eye;
  threshold 0 drawing the cloud as it is, a huge one only the coarsest level;
  the projected cell edge against float64.
"""
import ctypes as C

import numpy as np
import pytest

import fixtures as F
import oracle_lib as O


def _interleave(qx, qy, qz):
    code = 0
    for bit in range(9, -1, -1):
        code = (code << 3) | (((qx >> bit) & 1) << 2) | (((qy >> bit) & 1) << 1) | ((qz >> bit) & 1)
    return code


def test_morton_codes_interleave_ten_bits_an_axis_x_highest():
    rng = np.random.default_rng(1)
    units = list(rng.random((500, 3), dtype=np.float32)) + [np.float32([0, 0, 0]), np.float32([1, 1, 1]),
                                                            np.float32([-0.5, 2.0, 0.999]),
                                                            np.float32([0.5, 0.25, 0.75])]
    for u in units:
        q = [min(int(np.float32(min(max(v, 0.0), 1.0)) * np.float32(1024.0)), 1023) for v in u]
        assert O.lib.or_morton30(float(u[0]), float(u[1]), float(u[2])) == _interleave(*q), u


@pytest.fixture(scope="module")
def cloud():
    return F.random_cloud(5000, 41).decode(3)


@pytest.fixture(scope="module")
def lod(cloud):
    return O.lod_build(cloud.positions, cloud.shape, cloud.sh, cloud.rest_per_colour)


def test_morton_order_and_each_level_s_groups(cloud, lod):
    keys = lod["keys"].astype(np.uint64)
    assert np.all(np.diff(keys.astype(np.int64)) >= 0)
    np.testing.assert_array_equal(lod["positions"], cloud.positions[lod["order"]])
    # the sort is stable: equal codes keep index order
    same = np.diff(keys.astype(np.int64)) == 0
    assert np.all(np.diff(lod["order"].astype(np.int64))[same] > 0)
    n = cloud.count
    assert [lv["level"] for lv in lod["levels"]] == list(range(1, len(lod["levels"]) + 1))
    for lv in lod["levels"]:
        prefix = keys >> np.uint64(3 * (10 - lv["level"]))
        cells, starts = np.unique(prefix, return_index=True)
        np.testing.assert_array_equal(lv["cells"], cells.astype(np.uint32))
        np.testing.assert_array_equal(lv["starts"], starts.astype(np.uint32))
    finest = lod["levels"][-1]
    assert len(finest["cells"]) <= 0.5 * n
    deeper = np.unique(keys >> np.uint64(3 * (10 - finest["level"] - 1)))
    assert finest["level"] == 10 or len(deeper) > 0.5 * n


def _decoded(lod):
    """float64 parameters of the Morton-ordered splats."""
    n = lod["count"]
    pos = lod["positions"].astype(np.float64)
    shp = lod["shape"]
    h = lambda w, hi: (w >> 16 if hi else w & 0xFFFF).astype(np.uint16).view(np.float16).astype(np.float64)  # noqa: E731
    scales = np.exp(np.stack([h(shp[:, 1], 0), h(shp[:, 1], 1), h(shp[:, 2], 0)], axis=1))
    q = np.zeros((n, 4), np.float32)
    for i in range(n):
        qi = (C.c_float * 4)()
        O.lib.or_decode_quaternion(int(shp[i, 0]), qi)
        q[i] = qi[:]
    x, y, z, w = (q[:, k].astype(np.float64) for k in range(4))
    r = np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y),
                  2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x),
                  2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], axis=1).reshape(n, 3, 3)
    m = r * scales[:, None, :]
    cov = m @ m.transpose(0, 2, 1)
    area = scales.prod(axis=1) / scales.min(axis=1)
    weight = pos[:, 3] * area
    base = np.stack([h(shp[:, 2], 1), h(shp[:, 3], 0), h(shp[:, 3], 1)], axis=1)
    return pos, cov, weight, base


def test_leaf_moments_are_the_weighted_sums_of_their_splats(lod):
    pos, cov, weight, base = _decoded(lod)
    finest = lod["levels"][-1]
    starts = list(finest["starts"]) + [lod["count"]]
    m = finest["moments"].astype(np.float64)
    for g in range(0, len(finest["starts"]), 7):
        a, b = starts[g], starts[g + 1]
        w = weight[a:b]
        mu = pos[a:b, :3]
        want = [w.sum(), *(w @ mu)]
        s2 = np.einsum("i,ijk->jk", w, cov[a:b] + mu[:, :, None] * mu[:, None, :])
        want += [s2[0, 0], s2[0, 1], s2[0, 2], s2[1, 1], s2[1, 2], s2[2, 2], *(w @ base[a:b])]
        np.testing.assert_allclose(m[g, :13], want, rtol=2e-4, atol=1e-9)


def test_merged_moments_equal_the_moments_taken_at_that_level(lod):
    stride = 13 + 3 * lod["keep"]
    for lv in lod["levels"][:-1]:
        direct = np.zeros((len(lv["starts"]), stride), np.float32)
        O.lib.or_lod_leaf_moments(O.ptr(lod["positions"]), O.ptr(lod["shape"]), O.ptr(lod["sh"]), lod["count"],
                                  O.ptr(lv["starts"]), len(lv["starts"]), lod["keep"], lod["sh_words"], O.ptr(direct))
        np.testing.assert_allclose(lv["moments"], direct, rtol=2e-4, atol=1e-6)


def _gaussian(positions, shape):
    """float64 (mean, covariance, opacity) of packed Gaussians."""
    n = positions.shape[0]
    h = lambda w, hi: (w >> 16 if hi else w & 0xFFFF).astype(np.uint16).view(np.float16).astype(np.float64)  # noqa: E731
    s = np.exp(np.stack([h(shape[:, 1], 0), h(shape[:, 1], 1), h(shape[:, 2], 0)], axis=1))
    covs = np.zeros((n, 3, 3))
    for i in range(n):
        qi = (C.c_float * 4)()
        O.lib.or_decode_quaternion(int(shape[i, 0]), qi)
        x, y, z, w = (float(v) for v in qi[:])
        r = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                      [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                      [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
        covs[i] = r @ np.diag(s[i] ** 2) @ r.T
    return positions[:, :3].astype(np.float64), covs, positions[:, 3].astype(np.float64), s


def test_each_merged_gaussian_is_its_moments_turned_back_into_a_splat(lod):
    for lv in lod["levels"]:
        m = lv["moments"].astype(np.float64)
        w = m[:, 0]
        mu = m[:, 1:4] / w[:, None]
        s2 = np.stack([m[:, 4], m[:, 5], m[:, 6], m[:, 5], m[:, 7], m[:, 8], m[:, 6], m[:, 8], m[:, 9]],
                      axis=1).reshape(-1, 3, 3) / w[:, None, None]
        cov = s2 - mu[:, :, None] * mu[:, None, :]
        got_mu, got_cov, got_opacity, scales = _gaussian(lv["positions"], lv["shape"])
        np.testing.assert_allclose(got_mu, mu, rtol=1e-5, atol=1e-5)
        big = np.linalg.norm(cov, axis=(1, 2))
        err = np.linalg.norm(got_cov - cov, axis=(1, 2))
        # 1%: the 10-bit rotation and the f16 log scales; plus the float32
        # cancellation of S2 / W - mu mu^T, about 30 ulps of |mu|^2
        cancel = 2e-6 * (mu * mu).sum(axis=1)
        assert np.all(err <= 1e-2 * big + cancel), (lv["level"], float((err / (1e-2 * big + cancel)).max()))
        # the opacity is W over the area of the Gaussian's own two longest axes
        # (its stored scales: f16 logs, so to 1%)
        area = scales.prod(axis=1) / scales.min(axis=1)
        np.testing.assert_allclose(got_opacity, np.minimum(w / area, 0.99), rtol=1e-2)
        h = lambda v: (v >> 16).astype(np.uint16).view(np.float16).astype(np.float64)  # noqa: E731
        np.testing.assert_allclose(h(lv["shape"][:, 2]), m[:, 10] / w, rtol=2e-3, atol=2e-3)


def test_a_cell_holding_one_splat_merges_into_that_splat():
    """tests/lod/test_lod.cpp:107-153: six splats far apart, level 1 kept though it merges nothing."""
    b = F.CloudBuilder()
    rng = F.Lcg(0)
    at = [(-2, -2, -2), (2, -2, -2), (-2, 2, -2), (2, 2, 2), (-2, 2, 2), (2, -2, 2)]
    for p in at:
        opacity = rng.range(0.3, 0.95)
        sx, sy, sz = rng.range(0.1, 0.5), rng.range(0.1, 0.5), rng.range(0.1, 0.5)
        turn = [rng.range(-1, 1) for _ in range(4)]
        dc = [rng.range(0.1, 1.0) for _ in range(3)]
        b.add(p[0], p[1], p[2], opacity, sx, sy, sz, turn, dc, (0.2, -0.1, 0.3, 0.0, 0.1, -0.2, 0.1, 0.0, -0.3))
    sc = b.decode(3)
    lod = O.lod_build(sc.positions, sc.shape, sc.sh, 3, max_group_fraction=1.0, coarsest_level=1)
    level = lod["levels"][0]
    assert level["level"] == 1 and len(level["cells"]) == 6
    got_mu, got_cov, got_opacity, _ = _gaussian(level["positions"], level["shape"])
    want_mu, want_cov, want_opacity, _ = _gaussian(lod["positions"], lod["shape"])
    np.testing.assert_allclose(got_mu, want_mu, rtol=1e-6, atol=1e-6)
    for g in range(6):
        assert np.linalg.norm(got_cov[g] - want_cov[g]) <= 5e-3 * np.linalg.norm(want_cov[g])
    np.testing.assert_allclose(got_opacity, want_opacity, rtol=5e-3)
    np.testing.assert_allclose(level["sh"], lod["sh"], atol=1)   # the same halves, to a unit of the last place


def _flags(lod, eye, ppu, threshold, ortho=False):
    p = O.OrCutParams()
    p.lo[:] = [float(v) for v in lod["lo"]]
    p.extent = float(lod["extent"])
    p.eye[:] = [float(v) for v in eye]
    p.pixels_per_unit, p.threshold, p.orthographic = ppu, threshold, 1 if ortho else 0
    flags = []
    for k, lv in enumerate(lod["levels"]):
        sel = np.zeros(len(lv["cells"]), np.uint32)
        O.lib.or_lod_cut_groups(C.byref(p), O.ptr(lv["cells"]), len(lv["cells"]), lv["level"], 1 if k == 0 else 0,
                                O.ptr(sel))
        flags.append(sel)
    sel = np.zeros(lod["count"], np.uint32)
    fin = lod["levels"][-1]
    O.lib.or_lod_cut_splats(C.byref(p), O.ptr(lod["groups"]), lod["count"], O.ptr(fin["cells"]), fin["level"],
                            O.ptr(sel))
    return flags, sel, p


@pytest.mark.parametrize("eye,ppu,ortho", [((1.5, 1.0, 6.0), 300.0, False), ((0.1, 0.2, 0.3), 300.0, False),
                                           ((6.0, 4.0, 30.0), 800.0, False), ((0.0, 0.0, 50.0), 60.0, True)])
def test_the_cut_draws_every_place_exactly_once(lod, eye, ppu, ortho):
    for threshold in (0.0, 0.5, 2.0, 8.0, 30.0, 1e6):
        flags, splats, _ = _flags(lod, eye, ppu, threshold, ortho)
        covered = splats.astype(np.int64).copy()
        for lv, f in zip(lod["levels"], flags):
            covered += f[lv["group"]]
        assert np.all(covered == 1), (threshold, np.unique(covered, return_counts=True))
        if threshold == 0.0:
            assert splats.all() and not any(f.any() for f in flags)
        inside = np.all(np.asarray(eye) >= lod["lo"]) and np.all(np.asarray(eye) <= lod["lo"] + lod["extent"])
        if threshold == 1e6 and not inside:   # a cell holding the eye is never small
            assert not splats.any() and flags[0].all() and not any(f.any() for f in flags[1:])


def test_projected_cell_edges_against_float64(lod):
    _, _, p = _flags(lod, (1.5, 1.0, 6.0), 300.0, 1.0)
    lo = np.array(lod["lo"], np.float64)
    eye = np.array([1.5, 1.0, 6.0])
    for lv in lod["levels"]:
        for cell in lv["cells"][::17]:
            r = lv["level"]
            edge = float(lod["extent"]) / 2 ** r
            idx = np.array([F_compact(int(cell) >> 2), F_compact(int(cell) >> 1), F_compact(int(cell))], np.float64)
            c0 = lo + idx * edge
            d = np.clip(eye, c0, c0 + edge) - eye
            want = edge * 300.0 / max(np.linalg.norm(d), 1e-6)
            got = O.lib.or_lod_projected_edge(C.byref(p), int(cell), r)
            assert abs(got - want) <= 1e-5 * want, (r, cell, got, want)


def F_compact(v):
    out = 0
    for bit in range(10):
        out |= ((v >> (3 * bit)) & 1) << bit
    return out
