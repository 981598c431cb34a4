"""The splat file readers, on the host, and the oracle's decode of the
encodings they produce.

- PLY: the reference loader test's file (tests/scene/test_loading.cpp:38-60,
  four degree-1 records) written here and read back: count, encoding and the
  records bit for bit; the fixtures tests/golden/splats/{tiny,sh3}.ply (the
  reference's tests/data/splats) against numpy's own parse of their binary
  body; a PLY of mixed property types converted value by value; each refusal
  with the reference's code (SplatReaders.cpp:50-171, PlyHeader.cpp).
- .splat: the reference test's record (test_loading.cpp:166-191) read back as
  14 floats, and the oracle's decode of it against that test's known answers.
"""
import ctypes as C
import math
import os
import struct

import numpy as np
import pytest

import fixtures as F
import oracle_lib as O
from lucabrtrender_amd import raster
from lucabrtrender_amd._abi import LrtError

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "splats")
INVALID, UNSUPPORTED, IO_FAILURE = 3, 4, 5

PLY_NAMES = (["x", "y", "z", "nx", "ny", "nz", "f_dc_0", "f_dc_1", "f_dc_2"] + [f"f_rest_{k}" for k in range(9)] +
             ["opacity", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3"])
NP_TYPES = {"float": "<f4", "float32": "<f4", "double": "<f8", "float64": "<f8", "uchar": "u1", "uint8": "u1",
            "char": "i1", "int8": "i1", "ushort": "<u2", "uint16": "<u2", "short": "<i2", "int16": "<i2",
            "uint": "<u4", "uint32": "<u4", "int": "<i4", "int32": "<i4"}


def _write_ply(path, names, rows, types=None, fmt="binary_little_endian", extra_header=""):
    types = types or ["float"] * len(names)
    header = f"ply\nformat {fmt} 1.0\n{extra_header}element vertex {len(rows)}\n"
    header += "".join(f"property {t} {n}\n" for t, n in zip(types, names)) + "end_header\n"
    dtype = np.dtype([(n, NP_TYPES[t]) for n, t in zip(names, types)])
    body = np.array([tuple(r) for r in rows], dtype=dtype).tobytes()
    with open(path, "wb") as f:
        f.write(header.encode() + body)


def _numpy_parse(path):
    """An independent parse: header lines to a numpy dtype, the body as records."""
    data = open(path, "rb").read()
    end = data.index(b"end_header\n") + len(b"end_header\n")
    lines = data[:end].decode().splitlines()
    count, fields = 0, []
    for line in lines:
        w = line.split()
        if w[:2] == ["element", "vertex"]:
            count = int(w[2])
        elif w and w[0] == "property":
            fields.append((w[2], NP_TYPES[w[1]]))
    body = np.frombuffer(data, dtype=np.dtype(fields), count=count, offset=end)
    with np.errstate(over="ignore"):   # a double past float range casts to inf, as in C
        cols = [body[n].astype(np.float32) for n, _ in fields]
    return np.stack(cols, axis=1), [n for n, _ in fields]


def _reference_records():
    """test_loading.cpp:91-116: first, faint, broken (NaN x), last."""
    def rec(x=1.0, y=2.0, z=3.0, dc=(0.0, 0.0, 0.0), opacity=0.0, rot=(2.0, 0.0, 0.0, 0.0)):
        rest = [0.1 * (k + 1) for k in range(9)]
        return [x, y, z, 0.0, 0.0, 0.0, *dc, *rest, opacity, 0.0, 0.6931472, -0.6931472, *rot]
    return [rec(), rec(opacity=-10.0), rec(x=float("nan")),
            rec(x=-4.0, y=5.0, z=-6.0, dc=(1.0 / 0.28209479, 0.0, 0.0), opacity=1.0986123, rot=(0.0, 0.0, 0.0, -3.0))]


def test_the_reference_loader_test_file_reads_back_bit_for_bit(tmp_path):
    path = str(tmp_path / "decode.ply")
    rows = _reference_records()
    _write_ply(path, PLY_NAMES, rows)
    raw = raster.read_splats(path)
    assert raw.count == 4
    e = raw.encoding
    assert e["floats_per_record"] == 26 and e["rest_per_colour"] == 3 and e["file_per_colour"] == 3
    assert (e["x"], e["y"], e["z"], e["dc0"], e["rest_base"], e["opacity"], e["scale0"]) == (0, 1, 2, 6, 9, 18, 19)
    assert (e["rot_w"], e["rot_x"], e["rot_y"], e["rot_z"]) == (22, 23, 24, 25)
    assert e["rest_colour_outer"] == 1 and e["opacity_mode"] == e["scale_mode"] == e["colour_mode"] == 0
    want = np.array(rows, np.float32)
    np.testing.assert_array_equal(raw.records.view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("name,count,rest", [("tiny.ply", 64, 0), ("sh3.ply", 64, 15)])
def test_the_reference_fixtures_read_as_numpy_parses_them(name, count, rest):
    path = os.path.join(GOLDEN, name)
    raw = raster.read_splats(path)
    ref, names = _numpy_parse(path)
    assert raw.count == count and raw.encoding["rest_per_colour"] == rest
    np.testing.assert_array_equal(raw.records.view(np.uint32), ref.view(np.uint32))
    e = raw.encoding
    for field, prop in (("x", "x"), ("opacity", "opacity"), ("scale2", "scale_2"), ("rot_w", "rot_0"),
                        ("dc1", "f_dc_1")):
        assert e[field] == names.index(prop)
    if rest:
        assert e["rest_base"] == names.index("f_rest_0") and e["file_per_colour"] == 15


def test_mixed_property_types_are_converted_value_by_value(tmp_path):
    path = str(tmp_path / "mixed.ply")
    names = ["x", "y", "z", "opacity", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3",
             "f_dc_0", "f_dc_1", "f_dc_2", "red"]
    types = ["double", "float", "short", "float", "float", "float", "float", "uchar", "char", "ushort", "int",
             "float", "float", "float", "uchar"]
    rows = [[1.25, -2.5, -7, 0.5, -1.0, -2.0, -3.0, 200, -5, 65000, -123456, 0.1, 0.2, 0.3, 9],
            [1e300, 3.0, 32767, 0.0, 0.0, 0.0, 0.0, 0, 127, 0, 2**31 - 1, 0.0, 0.0, 0.0, 255]]
    _write_ply(path, names, rows, types, extra_header="comment mixed types\n")
    raw = raster.read_splats(path)
    ref, _ = _numpy_parse(path)
    np.testing.assert_array_equal(raw.records.view(np.uint32), ref.view(np.uint32))
    assert math.isinf(float(raw.records[1, 0]))   # a double past float range becomes inf, as a C cast does


def _refusal(path):
    with pytest.raises(LrtError) as err:
        raster.read_splats(str(path))
    return err.value.code


def test_refusals_carry_the_reference_codes(tmp_path):
    rows = _reference_records()[:1]
    _write_ply(tmp_path / "a.ply", PLY_NAMES, rows, fmt="ascii")
    assert _refusal(tmp_path / "a.ply") == UNSUPPORTED
    _write_ply(tmp_path / "b.ply", PLY_NAMES, rows, fmt="binary_big_endian")
    assert _refusal(tmp_path / "b.ply") == UNSUPPORTED
    _write_ply(tmp_path / "c.ply", [n for n in PLY_NAMES if n != "scale_1"], [r[:20] + r[21:] for r in rows])
    assert _refusal(tmp_path / "c.ply") == INVALID
    _write_ply(tmp_path / "d.ply", PLY_NAMES, rows)
    data = open(tmp_path / "d.ply", "rb").read()
    open(tmp_path / "d.ply", "wb").write(data[:-8])
    assert _refusal(tmp_path / "d.ply") == IO_FAILURE
    (tmp_path / "e.ply").write_bytes(b"ply\nformat binary_little_endian 1.0\nelement vertex 1\nproperty float x\n")
    assert _refusal(tmp_path / "e.ply") == INVALID                    # no end_header
    (tmp_path / "f.ply").write_bytes(b"not a ply\n")
    assert _refusal(tmp_path / "f.ply") == INVALID
    (tmp_path / "g.ply").write_bytes(b"ply\nformat binary_little_endian 1.0\nelement vertex 1\n"
                                     b"property list uchar int idx\nproperty float x\nend_header\n")
    assert _refusal(tmp_path / "g.ply") == UNSUPPORTED               # a list in the vertex element
    gap = [n for n in PLY_NAMES if n != "f_rest_4"] + ["f_rest_4"]    # rest present but out of order
    _write_ply(tmp_path / "h.ply", gap, [r[:13] + r[14:] + [r[13]] for r in rows])
    assert _refusal(tmp_path / "h.ply") == UNSUPPORTED
    header = (b"ply\nformat binary_little_endian 1.0\nelement face 1\nproperty list uchar int vertex_indices\n"
              b"element vertex 1\nproperty float x\nend_header\n")
    (tmp_path / "i.ply").write_bytes(header + b"\x00" * 16)
    assert _refusal(tmp_path / "i.ply") == UNSUPPORTED               # a list element before the vertices
    (tmp_path / "j.splat").write_bytes(b"\x00" * 33)
    assert _refusal(tmp_path / "j.splat") == INVALID
    (tmp_path / "k.spz").write_bytes(b"\x00" * 32)
    assert _refusal(tmp_path / "k.spz") == IO_FAILURE                  # not a gzip stream
    for name in ("l.sog", "m.xyz"):
        (tmp_path / name).write_bytes(b"\x00" * 32)
        assert _refusal(tmp_path / name) == UNSUPPORTED
    assert _refusal(tmp_path / "missing.ply") == IO_FAILURE


def test_a_header_whose_sizes_wrap_is_refused(tmp_path):
    """Element counts come from the header: an element before the vertices whose
    size puts the vertex offset just under 2^64 (so offset + vertex bytes wraps
    to a small number), and one whose stride * count wraps, are both refused
    with IoFailure instead of reading outside the file."""
    props = "".join(f"property float {n}\n" for n in PLY_NAMES)
    body = b"\x00" * (4 * len(PLY_NAMES) + 64)
    probe = f"ply\nformat binary_little_endian 1.0\nelement junk {{}}\nproperty uchar u\nelement vertex 1\n{props}end_header\n"
    digits = len(str(2 ** 64 - 1))
    hdr = len(probe.format("9" * digits))
    count = 2 ** 64 - 16 - hdr
    assert len(str(count)) == digits
    (tmp_path / "wrap_add.ply").write_bytes(probe.format(count).encode() + body)
    assert _refusal(tmp_path / "wrap_add.ply") == IO_FAILURE
    junk = "".join(f"property double d{k}\n" for k in range(8))     # a 64-byte stride
    header = (f"ply\nformat binary_little_endian 1.0\nelement junk {2 ** 62 + 1}\n{junk}element vertex 1\n"
              f"{props}end_header\n")
    (tmp_path / "wrap_mul.ply").write_bytes(header.encode() + body)
    assert _refusal(tmp_path / "wrap_mul.ply") == IO_FAILURE


def _splat_record():
    """test_loading.cpp:168-172."""
    return struct.pack("<6f8B", 0.5, -0.5, 2.0, 0.25, 0.5, 1.0, 255, 0, 51, 255, 255, 128, 128, 128)


def test_a_dot_splat_record_reads_as_fourteen_floats(tmp_path):
    path = tmp_path / "decode.splat"
    path.write_bytes(_splat_record() * 3)
    raw = raster.read_splats(str(path))
    assert raw.count == 3
    want = np.array([0.5, -0.5, 2.0, 0.25, 0.5, 1.0, 255, 0, 51, 255, 255, 128, 128, 128], np.float32)
    np.testing.assert_array_equal(raw.records, np.tile(want, (3, 1)))
    e = raw.encoding
    assert (e["opacity_mode"], e["scale_mode"], e["colour_mode"], e["rotation_mode"]) == (2, 1, 2, 1)
    assert (e["dc0"], e["opacity"], e["rot_w"], e["scale0"]) == (6, 9, 10, 3)


def _halves(word):
    return O.lib.or_half_to_float(int(word) & 0xFFFF), O.lib.or_half_to_float(int(word) >> 16)


def _quat(word):
    q = (C.c_float * 4)()
    O.lib.or_decode_quaternion(int(word), q)
    return np.array(q[:])


def test_the_oracle_decodes_the_dot_splat_known_answers(tmp_path):
    """test_loading.cpp:184-190: opacity 255 -> 1, scale 0.25 and 1 (linear,
    stored as f16 logs), rotation bytes (255, 128, 128, 128) -> w ~ 1, colour
    bytes 255 and 51 -> 1.0 and 0.2."""
    path = tmp_path / "decode.splat"
    path.write_bytes(_splat_record())
    raw = raster.read_splats(str(path))
    scene = F.decode_raw(raw, keep=3)
    assert scene.count == 1
    p, s = scene.positions[0], scene.shape[0]
    assert p[3] == 1.0
    l0, l1 = _halves(s[1])
    l2, r = _halves(s[2])
    g, b = _halves(s[3])
    assert math.exp(l0) == pytest.approx(0.25, rel=2e-3) and math.exp(l2) == pytest.approx(1.0, rel=2e-3)
    assert _quat(s[0])[3] == pytest.approx(1.0, abs=1e-2)
    assert r == pytest.approx(1.0, abs=1e-3) and b == pytest.approx(0.2, abs=1e-3) and g == 0.0


def test_raw_records_from_memory_and_their_checks():
    rows = np.array(_reference_records(), np.float32)
    enc = F.ply_encoding(rest_per_colour=3, file_per_colour=3, floats=26, rest_base=9, opacity=18, scale0=19,
                         rot_w=22, dc0=6)
    raw = raster.raw_splats(rows, enc)
    assert raw.count == 4 and raw.encoding == {**{k: 0 for k in raw.encoding}, **enc}
    np.testing.assert_array_equal(raw.records.view(np.uint32), rows.view(np.uint32))
    with pytest.raises(LrtError) as err:
        raster.raw_splats(rows, {**enc, "scale_mode": 7})
    assert err.value.code == INVALID
    with pytest.raises(LrtError) as err:
        raster.raw_splats(rows, {**enc, "rest_per_colour": 4})
    assert err.value.code == INVALID
