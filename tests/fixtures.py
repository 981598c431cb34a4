"""The reference's synthetic test clouds, rebuilt in Python.

tests/render/SplatFixtures.h: CloudBuilder writes 3DGS-encoded records (logit
opacity, log scales, (w, x, y, z) rotation, SH DC, nine rest coefficients
colour-outer) and randomCloud(count, seed) draws them from its LCG, each value
into a named variable in a fixed order. The float arithmetic is done in
float32 and exp/log come from the C library, as the C++ fixture computes them.
The records are then decoded by the oracle's restatement of the loader
(splat_validate + splat_decode) into the packed layout both renderers read.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

import oracle_lib as O
from lucabrtrender_amd.scenes import HostScene

_libm = C.CDLL("libm.so.6")
_libm.expf.restype = C.c_float
_libm.expf.argtypes = [C.c_float]
_libm.logf.restype = C.c_float
_libm.logf.argtypes = [C.c_float]
f32 = np.float32


def expf(v) -> np.float32:
    return f32(_libm.expf(float(v)))


def logf(v) -> np.float32:
    return f32(_libm.logf(float(v)))


class Lcg:
    """SplatFixtures.h:48-55."""

    def __init__(self, seed: int = 0):
        self.state = 0x853c49e6748fea9b ^ seed

    def next(self) -> np.float32:
        self.state = (self.state * 6364136223846793005 + 1442695040888963407) & 0xFFFFFFFFFFFFFFFF
        return f32((self.state >> 40) & 0xFFFFFF) / f32(16777216.0)

    def range(self, lo, hi) -> np.float32:
        return f32(lo) + (f32(hi) - f32(lo)) * self.next()


SH0 = f32(0.28209479)

# CloudBuilder's encoding (SplatFixtures.h:19-28): 23 floats a record.
ENCODING = dict(floats_per_record=23, x=0, y=1, z=2, opacity=3, scale0=4, scale1=5, scale2=6, rot_w=7, rot_x=8,
                rot_y=9, rot_z=10, dc0=11, dc1=12, dc2=13, rest_base=14, file_per_colour=3, rest_colour_outer=1)


class CloudBuilder:
    """SplatFixtures.h:16-46."""

    def __init__(self):
        self.records = []

    def add(self, x, y, z, opacity, sx, sy, sz, wxyz, colour, rest=(0.0,) * 9):
        a = f32(opacity)
        logit = logf(a / (f32(1.0) - a))
        r = [f32(x), f32(y), f32(z), logit, logf(sx), logf(sy), logf(sz), f32(wxyz[0]), f32(wxyz[1]),
             f32(wxyz[2]), f32(wxyz[3]),
             (f32(colour[0]) - f32(0.5)) / SH0, (f32(colour[1]) - f32(0.5)) / SH0, (f32(colour[2]) - f32(0.5)) / SH0]
        r += [f32(v) for v in rest]
        self.records.append(r)

    def array(self) -> np.ndarray:
        return np.array(self.records, dtype=np.float32).reshape(-1, 23)

    def decode(self, keep: int = 3) -> HostScene:
        return decode_records(self.array(), keep)


def decode_records(records: np.ndarray, keep: int = 3) -> HostScene:
    """The oracle's restatement of validate + decode (CloudLoader::upload)."""
    records = np.ascontiguousarray(records, dtype=np.float32)
    n = records.shape[0]
    e = O.OrEncoding(**ENCODING)
    words = 1 if keep == 0 else (keep * 3 + 1) // 2
    pos = np.zeros((n, 4), np.float32)
    shape = np.zeros((n, 4), np.uint32)
    sh = np.zeros((n, words), np.uint32)
    kept = O.lib.or_decode_records(O.ptr(records), n, C.byref(e), keep, O.ptr(pos), O.ptr(shape), O.ptr(sh), words)
    return HostScene(pos[:kept].copy(), shape[:kept].copy(), sh[:kept].copy() if keep else None, keep)


def decode_raw(raw, keep: int = 3) -> HostScene:
    """The oracle's decode of a library RawSplats (a file the readers read), with
    the rest coefficients kept up to `keep` a colour, as the GPU decode keeps them."""
    e = raw.encoding
    keep = min(keep, e["rest_per_colour"])
    enc = O.OrEncoding(**{k: v for k, v in e.items() if k != "rest_per_colour"})
    if not e["rest_colour_outer"]:
        enc.file_per_colour = keep
    records = np.ascontiguousarray(raw.records, dtype=np.float32)
    n = records.shape[0]
    words = 1 if keep == 0 else (keep * 3 + 1) // 2
    pos = np.zeros((n, 4), np.float32)
    shape = np.zeros((n, 4), np.uint32)
    sh = np.zeros((n, words), np.uint32)
    kept = O.lib.or_decode_records(O.ptr(records), n, C.byref(enc), keep, O.ptr(pos), O.ptr(shape), O.ptr(sh),
                                   words)
    return HostScene(pos[:kept].copy(), shape[:kept].copy(), sh[:kept].copy() if keep else None, keep)


def ply_encoding(rest_per_colour=0, file_per_colour=0, floats=14, rest_base=0, opacity=3, scale0=4, rot_w=7,
                 dc0=11, x=0):
    """A 3DGS-PLY lrt_splat_encoding: fields consecutive from each given start
    (position x y z, scales, rotation w x y z, dc), modes all 0."""
    return dict(floats_per_record=floats, x=x, y=x + 1, z=x + 2, opacity=opacity, scale0=scale0, scale1=scale0 + 1,
                scale2=scale0 + 2, rot_w=rot_w, rot_x=rot_w + 1, rot_y=rot_w + 2, rot_z=rot_w + 3, dc0=dc0,
                dc1=dc0 + 1, dc2=dc0 + 2, rest_base=rest_base, rest_per_colour=rest_per_colour,
                file_per_colour=file_per_colour, rest_colour_outer=1)


def random_cloud(count: int, seed: int, min_size=0.004, max_size=0.6, min_opacity=0.02,
                 max_opacity=0.99) -> CloudBuilder:
    """SplatFixtures.h:57-85, draw for draw."""
    b = CloudBuilder()
    rng = Lcg(seed)
    lo, hi = logf(min_size), logf(max_size)
    for _ in range(count):
        size = expf(rng.range(lo, hi))
        stretch = rng.range(0.05, 1.0)
        rest = [rng.range(-0.3, 0.3) for _ in range(9)]
        x = rng.range(-2.0, 2.0)
        y = rng.range(-1.5, 1.5)
        z = rng.range(-2.0, 2.0)
        opacity = rng.range(min_opacity, max_opacity)
        thin = rng.range(0.1, 1.0)
        turn = [rng.range(-1, 1) for _ in range(4)]
        dc = [rng.range(0.0, 1.2) for _ in range(3)]
        b.add(x, y, z, opacity, size, size * stretch, size * thin, turn, dc, rest)
    return b


class PointLcg:
    """tests/render/test_points.cpp:22-29 (its own starting state)."""

    def __init__(self, seed: int = 0):
        self.state = 0x2545F4914F6CDD1D ^ seed

    def next(self) -> np.float32:
        self.state = (self.state * 6364136223846793005 + 1442695040888963407) & 0xFFFFFFFFFFFFFFFF
        return f32((self.state >> 40) & 0xFFFFFF) / f32(16777216.0)

    def range(self, lo, hi) -> np.float32:
        return f32(lo) + (f32(hi) - f32(lo)) * self.next()


def pack_point_colours(rgb: np.ndarray) -> np.ndarray:
    """points_decode.slang:39-40: f16 r | g << 16, b | 1.0 << 16 (numpy's
    float16 conversion rounds to nearest even, as the reference's f32tof16)."""
    rgb = np.asarray(rgb, np.float32).reshape(-1, 3)
    h = np.concatenate([rgb, np.ones((rgb.shape[0], 1), np.float32)], axis=1).astype(np.float16).view(np.uint16)
    h = h.astype(np.uint32)
    return np.ascontiguousarray(np.stack([h[:, 0] | (h[:, 1] << 16), h[:, 2] | (h[:, 3] << 16)], axis=1))


def random_points(count: int, seed: int):
    """randomPoints (tests/render/test_points.cpp:31-45), draw for draw: positions
    float32 [N, 4] (x, y, z, 1) and packed linear colours uint32 [N, 2]."""
    rng = PointLcg(seed)
    xyz, rgb = [], []
    for _ in range(count):
        x, y, z = rng.range(-2, 2), rng.range(-1.5, 1.5), rng.range(-2, 2)
        r, g, b = rng.range(0, 1), rng.range(0, 1), rng.range(0, 1)
        xyz.append((x, y, z))
        rgb.append((r, g, b))
    positions = np.concatenate([np.array(xyz, np.float32), np.ones((count, 1), np.float32)], axis=1)
    return np.ascontiguousarray(positions), pack_point_colours(np.array(rgb, np.float32))


# The fog's encoding (tests/render/test_points.cpp:150-157): 14 floats, linear
# opacity, scales and colour.
FOG_ENCODING = dict(floats_per_record=14, x=0, y=1, z=2, opacity=3, scale0=4, scale1=5, scale2=6, rot_w=7, rot_x=8,
                    rot_y=9, rot_z=10, dc0=11, dc1=12, dc2=13, rest_base=0, file_per_colour=0, rest_colour_outer=1,
                    opacity_mode=1, scale_mode=1, colour_mode=1)


def fog_records() -> np.ndarray:
    """The 40 large translucent splats of test_points.cpp:158-165, draw for draw."""
    rng = PointLcg(0)
    rows = []
    for _ in range(40):
        x, y, z = rng.range(-2, 2), rng.range(-1, 1), rng.range(-2, 3)
        a = rng.range(0.1, 0.6)
        s0, s1, s2 = rng.range(0.1, 0.5), rng.range(0.1, 0.5), rng.range(0.1, 0.5)
        r, g, b = rng.range(0, 1), rng.range(0, 1), rng.range(0, 1)
        rows.append([x, y, z, a, s0, s1, s2, 1, 0, 0, 0, r, g, b])
    return np.array(rows, np.float32)


def decode_encoded(records: np.ndarray, encoding: dict, keep: int = 0) -> HostScene:
    """The oracle's decode of records in any encoding (OrEncoding's fields)."""
    records = np.ascontiguousarray(records, dtype=np.float32)
    n = records.shape[0]
    enc = O.OrEncoding(**encoding)
    words = 1 if keep == 0 else (keep * 3 + 1) // 2
    pos = np.zeros((n, 4), np.float32)
    shape = np.zeros((n, 4), np.uint32)
    sh = np.zeros((n, words), np.uint32)
    kept = O.lib.or_decode_records(O.ptr(records), n, C.byref(enc), keep, O.ptr(pos), O.ptr(shape), O.ptr(sh), words)
    return HostScene(pos[:kept].copy(), shape[:kept].copy(), sh[:kept].copy() if keep else None, keep)
