"""The HIP scan and onesweep sort through the C ABI, element for element
against the closed form and numpy's stable sort, on the reference's own test
inputs (tests/gpu/test_algo.cpp:72-109, 347-379)."""
import numpy as np
import pytest

import patterns as PT

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rast():
    from lucabrtrender_amd import raster
    return raster.TileRasterizer()


def _dev(a):
    import torch
    return torch.from_numpy(a.view(np.int64) if a.dtype == np.uint64 else a.view(np.int32)).cuda()


@pytest.mark.parametrize("count", PT.SCAN_SIZES)
def test_prefix_sum_of_every_size(rast, count):
    import torch
    c = PT.counts_generate(count)
    inp = _dev(c)
    out = torch.zeros_like(inp)
    total = torch.zeros(1, dtype=torch.int32, device="cuda")
    rast.prefix_sum(inp, out, total)
    torch.cuda.synchronize()
    want = np.concatenate([[0], np.cumsum(c.astype(np.uint64))[:-1]]).astype(np.uint32)
    assert np.array_equal(out.cpu().numpy().view(np.uint32), want)
    assert int(total.cpu().numpy().view(np.uint32)[0]) == int(c.astype(np.uint64).sum())


def test_an_empty_prefix_sum_sets_the_total_to_zero(rast):
    import torch
    inp = torch.ones(4, dtype=torch.int32, device="cuda")
    out = torch.full((4,), 7, dtype=torch.int32, device="cuda")
    total = torch.full((1,), 9, dtype=torch.int32, device="cuda")
    from lucabrtrender_amd import _abi
    assert _abi.lib.lrt_prefix_sum(rast._ctx, inp.data_ptr(), out.data_ptr(), total.data_ptr(), 0) == _abi.LRT_OK
    torch.cuda.synchronize()
    assert int(total.cpu()[0]) == 0 and out.cpu().tolist() == [7, 7, 7, 7]


def test_prefix_sum_on_unaligned_views(rast):
    import torch
    c = PT.counts_generate(50001, seed=9)
    base = torch.zeros(50003, dtype=torch.int32, device="cuda")
    base[1:50002] = _dev(c)
    out = torch.zeros(50003, dtype=torch.int32, device="cuda")
    total = torch.zeros(1, dtype=torch.int32, device="cuda")
    rast.prefix_sum(base[1:50002], out[2:50003], total)
    torch.cuda.synchronize()
    want = np.concatenate([[0], np.cumsum(c.astype(np.uint64))[:-1]]).astype(np.uint32)
    assert np.array_equal(out[2:50003].cpu().numpy().view(np.uint32), want)


@pytest.mark.parametrize("count,key_bits", PT.SORT_CASES + [(3_000_001, 37), (1_000_000, 8)])
@pytest.mark.parametrize("mode", range(5))
def test_radix_sort_orders_keeps_ties_stable_and_loses_nothing(rast, count, key_bits, mode):
    import torch
    keys, values = PT.sort_generate(count, mode, key_bits)
    want_k, want_v = PT.stable_sorted(keys, values, key_bits)
    k, v = _dev(keys), _dev(values)
    k2, v2 = torch.empty_like(k), torch.empty_like(v)
    in_scratch = rast.radix_sort(k, v, k2, v2, key_bits)
    torch.cuda.synchronize()
    rk, rv = (k2, v2) if in_scratch else (k, v)
    assert np.array_equal(rk.cpu().numpy().view(np.uint64), want_k)
    assert np.array_equal(rv.cpu().numpy().view(np.uint32), want_v)


@pytest.mark.parametrize("count,key_bits,radix_bits", [(777, 18, 9), (1_000_003, 27, 9), (100_000, 19, 10),
                                                      (200_001, 30, 10), (3_000_001, 37, 10), (50_000, 64, 9),
                                                      (2_000_000, 33, 9), (2_000_000, 33, 8)])
@pytest.mark.parametrize("mode", range(5))
def test_wider_digits_sort_the_same(rast, count, key_bits, radix_bits, mode):
    """Digits of 9 and 10 bits (and 33 bits as 7 + 7 + 7 + 6 + 6): the same
    stable order as numpy's, in ceil(key bits / digit bits) passes."""
    import torch
    keys, values = PT.sort_generate(count, mode, key_bits)
    want_k, want_v = PT.stable_sorted(keys, values, key_bits)
    k, v = _dev(keys), _dev(values)
    k2, v2 = torch.empty_like(k), torch.empty_like(v)
    in_scratch = rast.radix_sort(k, v, k2, v2, key_bits, radix_bits)
    torch.cuda.synchronize()
    assert in_scratch == (-(-key_bits // radix_bits) % 2 == 1)
    rk, rv = (k2, v2) if in_scratch else (k, v)
    assert np.array_equal(rk.cpu().numpy().view(np.uint64), want_k)
    assert np.array_equal(rv.cpu().numpy().view(np.uint32), want_v)


def test_sorting_ten_million_pairs_timed(rast):
    import torch
    keys, values = PT.sort_generate(10_000_000, 0, 32)
    k, v = _dev(keys), _dev(values)
    k2, v2 = torch.empty_like(k), torch.empty_like(v)
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    best = 1e9
    for _ in range(3):
        kk, vv = k.clone(), v.clone()
        torch.cuda.synchronize()
        start.record()
        in_scratch = rast.radix_sort(kk, vv, k2, v2, 32)
        stop.record()
        torch.cuda.synchronize()
        best = min(best, start.elapsed_time(stop))
    rk = (k2 if in_scratch else kk).cpu().numpy().view(np.uint64)
    assert np.all(np.diff(rk.astype(np.int64)) >= 0)
    print(f"radix sort, 10M pairs, 32-bit keys: {best:.3f} ms ({10e6 / best / 1e6:.2f} Gkeys/s)")
