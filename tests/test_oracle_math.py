"""Pins the oracle's own arithmetic: exp/log against libm in double, fp16
conversions against numpy on every half and on random floats, the sRGB curve
against its formula in double."""
import math

import numpy as np

import oracle_lib as O


def _ulps(got: np.ndarray, want: np.ndarray) -> np.ndarray:
    want32 = want.astype(np.float32)
    return np.abs(got.astype(np.float64) - want) / np.spacing(np.abs(want32)).astype(np.float64)


def test_expf_within_two_ulps_of_double_exp():
    rng = np.random.default_rng(1)
    xs = np.concatenate([np.linspace(-103.9, 88.7, 20001), rng.uniform(-20, 20, 20000),
                         rng.uniform(-1e-3, 1e-3, 2000), [-87.0, -88.0, -100.0, 0.5 * math.log(2)]]).astype(np.float32)
    got = np.array([O.lib.or_expf(float(x)) for x in xs], np.float32)
    want = np.exp(xs.astype(np.float64))
    ok = want < 3.4e38
    assert _ulps(got[ok], want[ok]).max() <= 2.0
    assert O.lib.or_expf(0.0) == 1.0
    assert math.isinf(O.lib.or_expf(89.0)) and O.lib.or_expf(-104.0) == 0.0
    assert math.isnan(O.lib.or_expf(float("nan")))
    # monotone on a dense grid
    grid = np.linspace(-5, 5, 4001).astype(np.float32)
    vals = np.array([O.lib.or_expf(float(x)) for x in grid])
    assert np.all(np.diff(vals) >= 0)


def test_logf_within_two_ulps_of_double_log():
    rng = np.random.default_rng(2)
    xs = np.concatenate([rng.uniform(1.0001, 255.0, 20000), np.exp(rng.uniform(-80, 80, 20000)),
                         [1.0, 1.0001, 2.0, 0.5, 1e-40, 3e-39, 255.0]]).astype(np.float32)
    xs = xs[xs > 0]
    got = np.array([O.lib.or_logf(float(x)) for x in xs], np.float32)
    want = np.log(xs.astype(np.float64))
    nz = np.abs(want) > 1e-30
    assert _ulps(got[nz], want[nz]).max() <= 2.0
    assert O.lib.or_logf(1.0) == 0.0
    assert math.isinf(O.lib.or_logf(0.0)) and math.isnan(O.lib.or_logf(-1.0))


def test_half_to_float_matches_numpy_on_every_half():
    h = np.arange(65536, dtype=np.uint16)
    want = h.view(np.float16).astype(np.float32)
    got = np.array([O.lib.or_half_to_float(int(v)) for v in h], np.float32)
    nan = np.isnan(want)
    assert np.array_equal(np.isnan(got), nan)
    assert np.array_equal(got[~nan].view(np.uint32), want[~nan].view(np.uint32))


def test_float_to_half_rounds_as_numpy():
    rng = np.random.default_rng(3)
    xs = np.concatenate([rng.normal(0, 1, 20000), rng.uniform(-70000, 70000, 5000),
                         rng.uniform(-1e-4, 1e-4, 5000), rng.uniform(-7e-8, 7e-8, 2000),
                         # ties at half precision: exactly between two halves
                         (np.arange(1, 2049, dtype=np.float64) + 0.5) * 2.0 ** -10,
                         [65504.0, 65519.0, 65520.0, -65520.0, 6.1e-5, 5.96e-8, 2.98e-8, 0.0, -0.0]]).astype(np.float32)
    with np.errstate(over="ignore"):   # past 65504 rounds to infinity, which is what is checked
        want = xs.astype(np.float16).view(np.uint16)
    got = np.array([O.lib.or_float_to_half(float(x)) for x in xs], np.uint16)
    assert np.array_equal(got, want)


def test_srgb_curve_matches_its_formula():
    cs = np.linspace(0.0, 1.6, 4001).astype(np.float32)
    got = np.array([O.lib.or_srgb_to_linear(float(c)) for c in cs], np.float64)
    c = cs.astype(np.float64)
    want = np.where(c <= 0.04045, c / 12.92, ((c + 0.055) / 1.055) ** 2.4)
    assert np.max(np.abs(got - want) / np.maximum(want, 1e-6)) < 2e-6
    assert O.lib.or_srgb_to_linear(-1.0) == 0.0
