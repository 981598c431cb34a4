"""Pins the oracle's scan and radix sort: the closed form (numpy cumsum) on the
reference's counts_generate input, and numpy's stable argsort on the five
sort_generate patterns at the reference's sizes and key widths."""
import numpy as np
import pytest

import oracle_lib as O
import patterns as PT


@pytest.mark.parametrize("count", PT.SCAN_SIZES)
def test_scan_is_the_closed_form(count):
    c = PT.counts_generate(count)
    out = np.zeros(count, np.uint32)
    total = O.lib.or_scan_exclusive(O.ptr(c), O.ptr(out), count)
    want = np.concatenate([[0], np.cumsum(c.astype(np.uint64))[:-1]])
    assert total == int(c.astype(np.uint64).sum())
    assert np.array_equal(out, want.astype(np.uint32))


@pytest.mark.parametrize("count,key_bits", PT.SORT_CASES)
@pytest.mark.parametrize("mode", range(5))
def test_radix_sort_is_a_stable_sort(count, key_bits, mode):
    keys, values = PT.sort_generate(count, mode, key_bits)
    want_k, want_v = PT.stable_sorted(keys, values, key_bits)
    k, v = keys.copy(), values.copy()
    O.lib.or_radix_sort(O.ptr(k), O.ptr(v), count, key_bits)
    assert np.array_equal(k, want_k)
    assert np.array_equal(v, want_v)


def test_radix_sort_ignores_bits_above_key_bits():
    keys = np.array([(1 << 40) | 5, 3, (1 << 41) | 3, 5], np.uint64)
    values = np.arange(4, dtype=np.uint32)
    O.lib.or_radix_sort(O.ptr(keys), O.ptr(values), 4, 8)
    assert list(values) == [1, 2, 0, 3]
