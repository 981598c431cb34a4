"""SPZ files (versions 2 and 3: one gzip stream) read on the host and decoded by
the oracle, against the reference loader test's known answers and against an
independent float64 parse of the same bytes.

- The reference's two cases (tests/scene/test_loading.cpp:227-316), written
  byte for byte as its writeSpz writes them: version 3 (smallest-three
  rotation, degree-1 harmonics, a transparent point dropped) and version 2
  (first-three rotation, no harmonics). Expected: right-up-back positions read
  right-down-front, opacity byte / 255, log scale byte / 16 - 10, colour
  through SPZ's DC scale 0.15, the quarter turn about y becoming one about -y,
  rest coefficients odd in y or z changing sign.
- Random files of every degree (0 to 4, degree 4 kept up to 3): the reader's
  records equal numpy's parse of the streams; the oracle's packed splats equal
  a float64 decode of the same bytes to the packing's precision.
- Refusals: version 1 (float16 positions) and version 4 (zstd) are
  Unsupported; a stream that is not gzip, a wrong magic, an unknown version and
  streams shorter than the header's point count are IoFailure.
"""
import ctypes as C
import gzip
import math
import struct

import numpy as np
import pytest

import fixtures as F
import oracle_lib as O
from lucabrtrender_amd import raster
from lucabrtrender_amd._abi import LrtError

MAGIC = 0x5053474E
UNSUPPORTED, IO_FAILURE = 4, 5
SH0 = 0.28209479177387814
FLIP = np.array([-1, -1, 1, -1, 1, 1, -1, 1, -1, 1, -1, -1, 1, -1, 1], np.float64)


def write_spz(path, version, points, sh_degree, fractional_bits, streams: bytes):
    """The reference test's writeSpz: a 16-byte header and the streams, gzipped together."""
    header = struct.pack("<IIIBBBB", MAGIC, version, points, sh_degree, fractional_bits, 0, 0)
    with gzip.open(path, "wb") as f:
        f.write(header + bytes(streams))


def int24(v):
    u = v & 0xFFFFFF
    return bytes([u & 0xFF, (u >> 8) & 0xFF, (u >> 16) & 0xFF])


def _halves(shape_words, k):
    """The f16 pair of a packed word as float64."""
    w = np.asarray(shape_words, np.uint32)
    return (w & 0xFFFF).astype(np.uint16).view(np.float16).astype(np.float64), \
           (w >> 16).astype(np.uint16).view(np.float16).astype(np.float64)


def _inspect(scene, i):
    """A packed splat as the reference test's inspect() lists it: position (3),
    opacity, scale (3), quaternion x y z w, base colour (3), rest coefficients."""
    pos = scene.positions[i]
    sh = scene.shape[i]
    q = (C.c_float * 4)()
    O.lib.or_decode_quaternion(int(sh[0]), q)
    l0, l1 = _halves(sh[1], 0)
    l2, b0 = _halves(sh[2], 0)
    b1, b2 = _halves(sh[3], 0)
    rest = []
    if scene.sh is not None:
        for w in scene.sh[i]:
            lo, hi = _halves(w, 0)
            rest += [float(lo), float(hi)]
    return [float(pos[0]), float(pos[1]), float(pos[2]), float(pos[3]), math.exp(l0), math.exp(l1), math.exp(l2),
            q[0], q[1], q[2], q[3], float(b0), float(b1), float(b2)] + rest


def test_version3_reference_case(tmp_path):
    s = bytearray()
    s += int24(4096) + int24(8192) + int24(-2048)            # (1, 2, -0.5) at 12 bits
    s += int24(0) + int24(0) + int24(0)
    s += bytes([191, 0])                                      # alphas: the second point is dropped
    s += bytes([255, 128, 0, 128, 128, 128])                  # colours
    s += bytes([160, 176, 144, 160, 160, 160])                # log scales 0, 1, -1
    comp = (3 << 30) | (511 << 10)                            # largest w; y = +1/sqrt2, x = z = 0
    s += struct.pack("<I", comp) * 2
    sh = bytes([192, 192, 192, 64, 64, 64, 192, 64, 128])    # rgb per basis: y +0.5, z -0.5, x mixed
    s += sh + sh
    path = str(tmp_path / "decode-v3.spz")
    write_spz(path, 3, 2, 1, 12, s)
    raw = raster.read_splats(path)
    e = raw.encoding
    assert raw.count == 2 and e["rest_per_colour"] == 3 and e["floats_per_record"] == 14 + 9
    assert (e["rotation_mode"], e["scale_mode"], e["colour_mode"], e["opacity_mode"], e["rest_mode"]) == (3, 2, 3, 2, 1)
    assert e["fractional_bits"] == 12 and e["flip_yz"] == 1 and e["rest_colour_outer"] == 0
    rec = raw.records
    assert rec[0, :3].tolist() == [4096.0, 8192.0, -2048.0] and rec[0, 10] == (comp & 0xFFFF) and rec[0, 11] == comp >> 16
    scene = F.decode_raw(raw, keep=15)
    assert scene.count == 1 and scene.rest_per_colour == 3
    v = _inspect(scene, 0)
    assert v[0:3] == [1.0, -2.0, 0.5]                          # right-up-back (1, 2, -0.5) read right-down-front
    assert v[3] == pytest.approx(191.0 / 255.0, abs=1e-6)
    assert v[4] == pytest.approx(1.0, rel=2e-3)
    assert v[5] == pytest.approx(math.exp(1.0), rel=2e-3)
    assert v[6] == pytest.approx(math.exp(-1.0), rel=2e-3)
    assert abs(v[8]) == pytest.approx(0.70710678, abs=3e-3) and abs(v[10]) == pytest.approx(0.70710678, abs=3e-3)
    assert v[8] * v[10] < 0.0                                  # a quarter turn about y becomes one about -y
    assert abs(v[7]) < 3e-3 and abs(v[9]) < 3e-3
    base = [0.5 + SH0 * ((b / 255.0 - 0.5) / 0.15) for b in (255, 128, 0)]
    assert v[11:14] == pytest.approx(base, abs=1e-3)
    assert v[14] == pytest.approx(-0.5, abs=1e-3) and v[15] == pytest.approx(-0.5, abs=1e-3)
    assert v[17] == pytest.approx(0.5, abs=1e-3) and v[19] == pytest.approx(0.5, abs=1e-3)


def test_version2_reference_case(tmp_path):
    s = bytearray()
    s += int24(-256) + int24(256) + int24(512)               # (-1, 1, 2) at 8 bits
    s += bytes([255]) + bytes([128, 128, 128]) + bytes([160, 160, 160])
    s += bytes([128, 128, 255])                               # x, y ~ 0, z = 1: a half turn about z
    path = str(tmp_path / "decode-v2.spz")
    write_spz(path, 2, 1, 0, 8, s)
    raw = raster.read_splats(path)
    assert raw.encoding["rotation_mode"] == 2 and raw.encoding["rest_per_colour"] == 0
    scene = F.decode_raw(raw, keep=15)
    assert scene.count == 1 and scene.rest_per_colour == 0
    v = _inspect(scene, 0)
    assert v[0:3] == [-1.0, -1.0, -2.0]
    assert v[3] == pytest.approx(1.0, abs=1e-6)
    assert abs(v[9]) == pytest.approx(1.0, abs=1e-2)


def _random_spz(rng, version, n, degree, bits):
    sh_dim = {0: 0, 1: 3, 2: 8, 3: 15, 4: 24}[degree]
    fixed = rng.integers(-(1 << 23), 1 << 23, (n, 3))
    alphas = rng.integers(0, 256, n).astype(np.uint8)
    colours = rng.integers(0, 256, (n, 3)).astype(np.uint8)
    scales = rng.integers(0, 256, (n, 3)).astype(np.uint8)
    rot = rng.integers(0, 256, (n, 4 if version >= 3 else 3)).astype(np.uint8)
    rest = rng.integers(0, 256, (n, sh_dim * 3)).astype(np.uint8)
    u = (fixed & 0xFFFFFF).astype(np.uint32)
    pos = np.stack([u & 0xFF, (u >> 8) & 0xFF, (u >> 16) & 0xFF], axis=2).astype(np.uint8).reshape(n, 9)
    streams = pos.tobytes() + alphas.tobytes() + colours.tobytes() + scales.tobytes() + rot.tobytes() + rest.tobytes()
    return streams, dict(fixed=fixed, alphas=alphas, colours=colours, scales=scales, rot=rot, rest=rest, sh_dim=sh_dim)


def _quaternion_f64(rot, version):
    """(x, y, z, w) in float64 from the rotation bytes, before the flip."""
    if version == 2:
        xyz = rot.astype(np.float64) / 127.5 - 1.0
        return np.append(xyz, math.sqrt(max(0.0, 1.0 - float(xyz @ xyz))))
    comp = int(rot[0]) | int(rot[1]) << 8 | int(rot[2]) << 16 | int(rot[3]) << 24
    largest = comp >> 30
    q = np.zeros(4)
    for i in (3, 2, 1, 0):
        if i == largest:
            continue
        m = math.sqrt(0.5) * (comp & 511) / 511.0
        q[i] = -m if (comp >> 9) & 1 else m
        comp >>= 10
    q[largest] = math.sqrt(max(0.0, 1.0 - float(q @ q)))
    return q


@pytest.mark.parametrize("version,degree,bits", [(3, 0, 12), (3, 1, 10), (3, 3, 12), (3, 4, 8), (2, 2, 16), (2, 3, 0)])
def test_random_files_against_a_float64_parse(tmp_path, version, degree, bits):
    rng = np.random.default_rng(version * 10 + degree)
    n = 2000
    streams, ref = _random_spz(rng, version, n, degree, bits)
    path = str(tmp_path / "r.spz")
    write_spz(path, version, n, degree, bits, streams)
    raw = raster.read_splats(path)
    rec = raw.records
    kept = min(ref["sh_dim"], 15)
    assert raw.count == n and rec.shape == (n, 14 + 3 * kept)
    np.testing.assert_array_equal(rec[:, 0:3], ref["fixed"].astype(np.float32))
    np.testing.assert_array_equal(rec[:, 3], ref["alphas"])
    np.testing.assert_array_equal(rec[:, 4:7], ref["scales"])
    np.testing.assert_array_equal(rec[:, 7:10], ref["colours"])
    if version == 3:
        r = ref["rot"].astype(np.uint32)
        np.testing.assert_array_equal(rec[:, 10], r[:, 0] | r[:, 1] << 8)
        np.testing.assert_array_equal(rec[:, 11], r[:, 2] | r[:, 3] << 8)
    else:
        np.testing.assert_array_equal(rec[:, 10:13], ref["rot"])
    np.testing.assert_array_equal(rec[:, 14:], ref["rest"][:, :3 * kept])
    scene = F.decode_raw(raw, keep=15)
    keep = ref["alphas"].astype(np.float64) / 255.0 >= 1.0 / 255.0
    assert scene.count == int(keep.sum())
    idx = np.nonzero(keep)[0]
    scale = 2.0 ** -bits
    want = ref["fixed"][idx].astype(np.float64) * scale * np.array([1.0, -1.0, -1.0])
    np.testing.assert_array_equal(scene.positions[:, :3].astype(np.float64), want)
    np.testing.assert_allclose(scene.positions[:, 3], ref["alphas"][idx] / 255.0, atol=1e-7)
    l0, l1 = _halves(scene.shape[:, 1], 0)
    l2, b0 = _halves(scene.shape[:, 2], 0)
    b1, b2 = _halves(scene.shape[:, 3], 0)
    np.testing.assert_allclose(np.stack([l0, l1, l2], 1), ref["scales"][idx] / 16.0 - 10.0, atol=8e-3)
    base = 0.5 + SH0 * ((ref["colours"][idx] / 255.0 - 0.5) / 0.15)
    np.testing.assert_allclose(np.stack([b0, b1, b2], 1), base, atol=2e-3)
    q = (C.c_float * 4)()
    for j, i in enumerate(idx[:300]):
        O.lib.or_decode_quaternion(int(scene.shape[j, 0]), q)
        got = np.array(q[:])
        w = _quaternion_f64(ref["rot"][i], version) * np.array([1.0, -1.0, -1.0, 1.0])
        w = w / np.linalg.norm(w) if np.linalg.norm(w) > 1e-8 else np.array([0.0, 0.0, 0.0, 1.0])
        assert min(np.abs(got - w).max(), np.abs(got + w).max()) < 3e-3, (i, got, w)
    if kept:
        halves = []
        for word in scene.sh.reshape(-1):
            halves += [word & 0xFFFF, word >> 16]
        vals = np.array(halves, np.uint32).astype(np.uint16).view(np.float16).astype(np.float64)
        vals = vals.reshape(scene.count, -1)[:, :3 * kept].reshape(scene.count, kept, 3)
        want_rest = ((ref["rest"][idx][:, :3 * kept].astype(np.float64) - 128.0) / 128.0).reshape(-1, kept, 3)
        want_rest *= FLIP[:kept][None, :, None]
        np.testing.assert_allclose(vals, want_rest, atol=1e-3)


def _refusal(path):
    with pytest.raises(LrtError) as err:
        raster.read_splats(str(path))
    return err.value.code


def test_refusals(tmp_path):
    one = int24(0) * 3 + bytes([255, 128, 128, 128, 160, 160, 160, 128, 128, 128, 128])
    write_spz(tmp_path / "v1.spz", 1, 1, 0, 12, one)
    assert _refusal(tmp_path / "v1.spz") == UNSUPPORTED              # float16 positions
    (tmp_path / "v4.spz").write_bytes(struct.pack("<IIIBBBBI", MAGIC, 4, 1, 0, 12, 0, 1, 32) + b"\0" * 12)
    assert _refusal(tmp_path / "v4.spz") == UNSUPPORTED              # zstd streams
    (tmp_path / "raw.spz").write_bytes(b"\x01" * 40)
    assert _refusal(tmp_path / "raw.spz") == IO_FAILURE              # not gzip
    with gzip.open(tmp_path / "magic.spz", "wb") as f:
        f.write(struct.pack("<IIIBBBB", 0x12345678, 3, 1, 0, 12, 0, 0) + one)
    assert _refusal(tmp_path / "magic.spz") == IO_FAILURE
    write_spz(tmp_path / "v9.spz", 9, 1, 0, 12, one)
    assert _refusal(tmp_path / "v9.spz") == IO_FAILURE
    write_spz(tmp_path / "short.spz", 3, 2, 0, 12, one)              # streams for one point, two declared
    assert _refusal(tmp_path / "short.spz") == IO_FAILURE
    write_spz(tmp_path / "none.spz", 3, 0, 0, 12, b"")
    assert _refusal(tmp_path / "none.spz") == IO_FAILURE
