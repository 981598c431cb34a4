"""The prefix sum that allocates nothing (lrt_prefix_sum_static), through the C
ABI: its output and total equal lrt_prefix_sum's and numpy's on two calls in a
row on one device (the second smaller, so the first call's scratch contents are
stale), at and around partition boundaries, and on an empty input."""
import ctypes as C

import numpy as np
import pytest
import torch

from lucabrtrender_amd import _abi, raster

pytestmark = pytest.mark.gpu


def _scan_both(rast, data):
    """(context output, context total, static output, static total) of one input."""
    n = data.size
    inp = torch.from_numpy((data if n else np.zeros(1, np.uint32)).view(np.int32)).cuda()
    out_ctx = torch.full((max(n, 1),), -1, dtype=torch.int32, device="cuda")
    out_static = torch.full((max(n, 1),), -7, dtype=torch.int32, device="cuda")
    tot_ctx = torch.full((1,), -1, dtype=torch.int32, device="cuda")
    tot_static = torch.full((1,), -7, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    assert _abi.lib.lrt_prefix_sum(rast._ctx, C.c_void_p(inp.data_ptr()), C.c_void_p(out_ctx.data_ptr()),
                                   C.c_void_p(tot_ctx.data_ptr()), n) == _abi.LRT_OK
    stream = torch.cuda.current_stream().cuda_stream
    torch.cuda.synchronize()
    code = _abi.lib.lrt_prefix_sum_static(C.c_void_p(inp.data_ptr()), C.c_void_p(out_static.data_ptr()),
                                          C.c_void_p(tot_static.data_ptr()), n, C.c_void_p(stream))
    assert code == _abi.LRT_OK, _abi.lib.lrt_last_error(None).decode()
    torch.cuda.synchronize()
    return (out_ctx[:n].cpu().numpy().view(np.uint32), int(tot_ctx.cpu().numpy().view(np.uint32)[0]),
            out_static[:n].cpu().numpy().view(np.uint32), int(tot_static.cpu().numpy().view(np.uint32)[0]))


@pytest.mark.parametrize("sizes", [(3_000_000, 1_000_003), (4096 * 5 + 1, 4096), (70_000, 0)])
def test_static_scan_equals_the_context_scan_on_two_calls(sizes):
    rast = raster.TileRasterizer()
    rng = np.random.default_rng(sizes[0] + sizes[1])
    for n in sizes:
        print(f"static prefix sum: {n} elements", flush=True)
        data = rng.integers(0, 300, n, dtype=np.uint32)
        data[rng.random(n) < 0.3] = 0
        out_ctx, tot_ctx, out_static, tot_static = _scan_both(rast, data)
        want = np.concatenate([[0], np.cumsum(data.astype(np.uint64))[:-1]]).astype(np.uint32) if n else data
        assert np.array_equal(out_static, out_ctx) and np.array_equal(out_static, want)
        assert tot_static == tot_ctx == int(data.astype(np.uint64).sum())


def test_static_scan_refuses_more_than_its_maximum():
    big = (1 << 28) + 1
    a = torch.zeros(4, dtype=torch.int32, device="cuda")
    b = torch.zeros(4, dtype=torch.int32, device="cuda")
    assert _abi.lib.lrt_prefix_sum_static(C.c_void_p(a.data_ptr()), C.c_void_p(b.data_ptr()),
                                          C.c_void_p(b.data_ptr()), big, None) == 3
    assert "exceeds" in _abi.lib.lrt_last_error(None).decode()
