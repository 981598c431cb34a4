"""Exact-arithmetic checks of shortcuts the kernels take for IEEE operations
(CPU only): each must give the correctly rounded result on its whole domain."""
from fractions import Fraction

import numpy as np


def _round_f32(x: Fraction) -> np.float32:
    """x rounded to the nearest float32, ties to even (x within float32 range)."""
    f = np.float32(float(x))
    cands = [np.nextafter(f, np.float32(-np.inf)), f, np.nextafter(f, np.float32(np.inf))]
    return min(cands, key=lambda c: (abs(Fraction(float(c)) - x), int(np.array(c).view(np.uint32)) & 1))


def _fma(a, b, c) -> np.float32:
    return _round_f32(Fraction(float(a)) * Fraction(float(b)) + Fraction(float(c)))


def test_div1023_is_the_correctly_rounded_quotient_for_every_10_bit_integer():
    """splat_math.h div1023: q = s * RN(1/1023), then q + RN(s - 1023 q) * RN(1/1023),
    both steps fused multiply-adds, equals RN(s / 1023) for s = 0 .. 1023 (the 10-bit
    fields of the smallest-three quaternion)."""
    r = np.float32(1.0) / np.float32(1023.0)
    for si in range(1024):
        s = np.float32(si)
        q = np.float32(s * r)
        e = _fma(-q, np.float32(1023.0), s)
        got = _fma(e, r, q)
        assert got == _round_f32(Fraction(si, 1023)), si


def test_the_plain_product_with_the_reciprocal_is_not_exact():
    """Without the correction step 24 of the 1,024 quotients are off by an ulp: the
    check above is not vacuous."""
    r = np.float32(1.0) / np.float32(1023.0)
    s = np.arange(1024, dtype=np.float32)
    assert int(np.sum(s * r != s / np.float32(1023.0))) == 24
