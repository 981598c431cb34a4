"""The GPU decode of raw splat records (lrt_cloud_decode) against the oracle's
restatement of splat_validate + splat_decode (pinned on the CPU by
tests/test_oracle_pack.py and tests/test_loader.py), bit for bit: positions,
packed shape words and harmonic words, the kept count (file order kept), and
the bounds; then frames drawn from a decoded file against frames drawn from
the oracle's packing of the same file.

Cases: the reference loader test's four PLY records (test_loading.cpp:91-150,
at degree 3 and capped at 0); its .splat record; the reference's fixtures
tiny.ply (degree 0) and sh3.ply (degree 3, capped at 1 as well); a 1.1M-record
PLY-encoded set with invalid records mixed in, large enough to be decoded in
two slices; and a set where nothing survives.
"""
import os

import numpy as np
import pytest
import torch

import fixtures as F
from lucabrtrender_amd import raster
from lucabrtrender_amd._abi import LrtError

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "splats")


@pytest.fixture(scope="module")
def rast():
    return raster.TileRasterizer()


def _check(rast, raw, max_degree):
    cloud, info = rast.decode_cloud(raw, max_degree)
    torch.cuda.synchronize()
    ref = F.decode_raw(raw, keep={0: 0, 1: 3, 2: 8}.get(max_degree, 15))
    print(f"  {raw.count} records, {info.count} kept, rest {info.rest_per_colour}, sh words {info.sh_words}",
          flush=True)
    assert info.declared == raw.count and info.count == ref.count == cloud.count
    assert info.rest_per_colour == ref.rest_per_colour
    np.testing.assert_array_equal(cloud.positions.cpu().numpy().view(np.uint32), ref.positions.view(np.uint32))
    np.testing.assert_array_equal(cloud.shape.cpu().numpy().view(np.uint32), ref.shape)
    if ref.rest_per_colour:
        np.testing.assert_array_equal(cloud.sh.cpu().numpy().view(np.uint32), ref.sh)
    else:
        assert cloud.sh is None
    xyz = ref.positions[:, :3]
    np.testing.assert_array_equal(np.array(info.bounds_min, np.float32), xyz.min(axis=0))
    np.testing.assert_array_equal(np.array(info.bounds_max, np.float32), xyz.max(axis=0))
    return cloud, ref


def _reference_ply(tmp_path):
    from test_loader import PLY_NAMES, _reference_records, _write_ply
    path = str(tmp_path / "decode.ply")
    _write_ply(path, PLY_NAMES, _reference_records())
    return raster.read_splats(path)


@pytest.mark.parametrize("max_degree", [3, 0])
def test_the_reference_loader_records_decode_as_the_oracle_does(rast, tmp_path, max_degree):
    raw = _reference_ply(tmp_path)
    cloud, _ = _check(rast, raw, max_degree)
    assert cloud.count == 2
    pos = cloud.positions.cpu().numpy()
    assert tuple(pos[0, :3]) == (1.0, 2.0, 3.0) and tuple(pos[1, :3]) == (-4.0, 5.0, -6.0)   # file order kept
    assert pos[0, 3] == pytest.approx(0.5, abs=1e-6) and pos[1, 3] == pytest.approx(0.75, abs=1e-5)


def test_a_dot_splat_record_decodes_as_the_oracle_does(rast, tmp_path):
    from test_loader import _splat_record
    path = tmp_path / "decode.splat"
    path.write_bytes(_splat_record() * 5)
    _check(rast, raster.read_splats(str(path)), 3)


@pytest.mark.parametrize("name,max_degree", [("tiny.ply", 3), ("sh3.ply", 3), ("sh3.ply", 1)])
def test_the_reference_fixtures_decode_and_draw_as_the_oracle_packs_them(rast, name, max_degree):
    raw = raster.read_splats(os.path.join(GOLDEN, name))
    cloud, ref = _check(rast, raw, max_degree)
    # The decoded cloud and the oracle-packed one draw the same frame, bit for bit.
    s = raster.RenderSettings(width=160, height=120, background=(0.1, 0.2, 0.3, 0.5))
    lo, hi = np.array(ref.positions[:, :3].min(0)), np.array(ref.positions[:, :3].max(0))
    centre = (lo + hi) / 2.0
    cam = raster.Camera.looking_at(tuple(centre + np.array([0.0, 0.0, 2.5 * float(np.max(hi - lo)) + 1.0])),
                                   tuple(centre))
    images = []
    for c in (cloud, ref.to_device()):
        t = raster.RenderTargets(160, 120)
        stats = rast.render(cam, [raster.SplatInstance(c)], s, t)
        torch.cuda.synchronize()
        images.append((t.colour.cpu().numpy(), stats.visible))
    assert images[0][1] > 0
    np.testing.assert_array_equal(images[0][0].view(np.uint32), images[1][0].view(np.uint32))


def test_a_million_records_decode_in_slices_as_the_oracle_does(rast):
    rng = np.random.default_rng(5)
    n, floats = 1_100_000, 62
    rec = np.zeros((n, floats), np.float32)
    rec[:, 0:3] = rng.uniform(-50, 50, (n, 3))
    rec[:, 3:6] = rng.normal(0, 1, (n, 3))                  # dc
    rec[:, 6:51] = rng.normal(0, 0.3, (n, 45))              # rest, colour-outer
    rec[:, 51] = rng.normal(0, 3, n)                        # opacity logit: some below 1/255
    rec[:, 52:55] = rng.normal(-4, 2, (n, 3))               # log scales
    rec[:, 55:59] = rng.normal(0, 1, (n, 4))                # rotation w x y z, unnormalised
    bad = rng.choice(n, 5000, replace=False)
    rec[bad[:2000], 0] = np.nan
    rec[bad[2000:4000], 53] = 70.0                          # |log scale| >= 60
    rec[bad[4000:], 55:59] = 0.0                            # zero rotation: identity
    enc = F.ply_encoding(rest_per_colour=15, file_per_colour=15, floats=floats, rest_base=6, opacity=51, scale0=52,
                         rot_w=55, dc0=3)
    raw = raster.raw_splats(rec, enc)
    cloud, ref = _check(rast, raw, 3)
    assert 0 < cloud.count < n


def test_nothing_surviving_is_an_invalid_argument(rast):
    rec = np.zeros((10, 14), np.float32)
    rec[:, 3] = -20.0   # opacity logit far below 1/255
    raw = raster.raw_splats(rec, F.ply_encoding())
    with pytest.raises(LrtError) as err:
        rast.decode_cloud(raw, 3)
    assert err.value.code == 3


@pytest.mark.parametrize("version,degree,bits", [(3, 1, 12), (3, 4, 10), (2, 3, 8), (2, 0, 16)])
def test_spz_files_decode_as_the_oracle_does(rast, tmp_path, version, degree, bits):
    """SPZ (version 2: first-three rotation; 3: smallest three), every degree
    path including degree 4 kept up to 3, fixed point at several widths,
    right-up-back flipped: the GPU decode against the oracle's, bit for bit."""
    from test_spz import _random_spz, write_spz
    rng = np.random.default_rng(100 + version * 10 + degree)
    n = 20_000
    streams, _ = _random_spz(rng, version, n, degree, bits)
    path = str(tmp_path / "r.spz")
    write_spz(path, version, n, degree, bits, streams)
    raw = raster.read_splats(path)
    print(f"SPZ v{version} degree {degree}, {bits} fractional bits", flush=True)
    for max_degree in (3, 1):
        _check(rast, raw, max_degree)
