"""Pins the oracle's packing and record decode: the known answers of the
reference's loader test (tests/scene/test_loading.cpp:91-165), quaternion
round trips, and agreement of the library's host packer with the oracle on the
fields both write from the same values."""
import ctypes as C
import math

import numpy as np

import fixtures as F
import oracle_lib as O
from lucabrtrender_amd import scenes


def _halves(word):
    return O.lib.or_half_to_float(word & 0xFFFF), O.lib.or_half_to_float(word >> 16)


def _quat(word):
    q = (C.c_float * 4)()
    O.lib.or_decode_quaternion(int(word), q)
    return np.array(q[:])


def _record(x=1.0, y=2.0, z=3.0, opacity=0.0, scale=(0.0, 0.6931472, -0.6931472), rot=(2.0, 0.0, 0.0, 0.0),
            dc=(0.0, 0.0, 0.0), rest=None):
    rest = rest if rest is not None else [0.1 * (k + 1) for k in range(9)]
    return [x, y, z, opacity, *scale, *rot, *dc, *rest]


def test_decode_known_answers_of_the_reference_loader_test():
    nan = float("nan")
    first = _record()
    faint = _record(opacity=-10.0)                    # opacity 4.5e-5 < 1/255: dropped
    broken = _record(x=nan)                           # NaN position: dropped
    last = _record(x=-4.0, y=5.0, z=-6.0, opacity=1.0986123, rot=(0.0, 0.0, 0.0, -3.0),
                   dc=(1.0 / 0.28209479, 0.0, 0.0))
    scene = F.decode_records(np.array([first, faint, broken, last], np.float32), keep=3)
    assert scene.count == 2                            # kept in file order
    p, s, sh = scene.positions, scene.shape, scene.sh
    assert tuple(p[0, :3]) == (1.0, 2.0, 3.0)
    assert abs(p[0, 3] - 0.5) <= 1e-6
    ls0, ls1 = _halves(int(s[0, 1]))
    ls2, base_r = _halves(int(s[0, 2]))
    assert abs(O.lib.or_expf(ls0) - 1.0) <= 2e-3
    assert abs(O.lib.or_expf(ls1) - 2.0) <= 2 * 2e-3
    assert abs(O.lib.or_expf(ls2) - 0.5) <= 0.5 * 2e-3
    assert abs(_quat(s[0, 0])[3] - 1.0) <= 1e-3        # w: the identity
    assert abs(base_r - 0.5) <= 1e-3
    # rgb per basis: basis 1 is (r .1, g .4, b .7), basis 2 starts (r .2, ...)
    h = [v for w in sh[0] for v in _halves(int(w))]
    assert np.allclose(h[:4], [0.1, 0.4, 0.7, 0.2], atol=1e-3)
    assert p[1, 0] == -4.0
    assert abs(p[1, 3] - 0.75) <= 1e-5
    assert abs(abs(_quat(s[1, 0])[2]) - 1.0) <= 1e-3   # |z| = 1
    assert abs(_halves(int(s[1, 2]))[1] - 1.5) <= 1e-3
    flat = F.decode_records(np.array([first, faint, broken, last], np.float32), keep=0)
    assert flat.count == 2 and flat.sh is None


def test_quaternion_round_trip_within_ten_bit_precision():
    rng = np.random.default_rng(5)
    worst = 0.0
    for _ in range(4000):
        q = rng.normal(size=4).astype(np.float32)
        word = O.lib.or_encode_quaternion((C.c_float * 4)(*q.tolist()))
        d = _quat(word)
        qn = q / np.linalg.norm(q)
        worst = max(worst, min(np.abs(d - qn).max(), np.abs(d + qn).max()))
        assert (word & 3) == int(np.argmax(np.abs(qn)))
    # three components on a 1023-step grid over [-sqrt(1/2), sqrt(1/2)], the fourth from unit length
    assert worst < 2.5e-3


def test_host_packer_agrees_with_the_oracle_on_shared_fields():
    rng = np.random.default_rng(6)
    n = 500
    fields = np.zeros((n, 14 + 45), np.float32)
    fields[:, 0:3] = rng.uniform(-3, 3, (n, 3))
    fields[:, 3] = rng.uniform(0.01, 1.0, n)
    fields[:, 4:7] = rng.normal(-4, 1, (n, 3))
    fields[:, 7:11] = rng.normal(size=(n, 4))
    fields[:, 11:14] = rng.uniform(0, 1, (n, 3))
    fields[:, 14:] = rng.normal(0, 0.3, (n, 45))
    packed = scenes.pack_splats(fields, 15)
    for i in range(0, n, 7):
        f = fields[i]
        assert packed.shape[i, 0] == O.lib.or_encode_quaternion((C.c_float * 4)(*f[7:11].tolist()))
        assert packed.shape[i, 1] == O.lib.or_pack_halves(float(f[4]), float(f[5]))
        assert packed.shape[i, 2] == O.lib.or_pack_halves(float(f[6]), float(f[11]))
        assert packed.shape[i, 3] == O.lib.or_pack_halves(float(f[12]), float(f[13]))
        want_sh = [O.lib.or_pack_halves(float(f[14 + 2 * k]), float(f[15 + 2 * k])) for k in range(22)]
        want_sh.append(O.lib.or_float_to_half(float(f[14 + 44])))
        assert list(packed.sh[i]) == want_sh
    assert np.array_equal(packed.positions[:, :3], fields[:, :3])


def test_random_cloud_is_the_reference_fixture_shape():
    b = F.random_cloud(3000, 7)
    recs = b.array()
    assert recs.shape == (3000, 23)
    scene = b.decode(3)
    # opacity >= 0.02 > 1/255 and finite log scales: nothing is dropped
    assert scene.count == 3000 and scene.sh.shape == (3000, 5)
    assert np.all(np.abs(scene.positions[:, 0]) <= 2.0) and np.all(np.abs(scene.positions[:, 1]) <= 1.5)
    assert math.isclose(float(np.mean(scene.positions[:, 3])), 0.505, abs_tol=0.03)
