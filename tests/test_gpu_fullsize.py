"""The sort/duplication stress frame at its full size (6,000,000 splats at
3840x2160, about 7e8 (tile, splat) pairs: more than the oracle can draw in a
test), checked by properties the sort must have whatever the inputs:

- the pairs are sorted by (tile, 24-bit depth key of the pair's slot), and
  pairs with equal keys are in slot order (the sort is stable over the
  duplication's slot-order output);
- every slot appears exactly as many times as the tiles it touches, and each
  pair's tile lies inside its slot's tile rectangle;
- each tile's range is exactly its run of pairs, and the runs cover every pair.
"""
import numpy as np
import pytest

from lucabrtrender_amd import _abi, raster, scenes

pytestmark = pytest.mark.gpu

CHUNK = 1 << 26


@pytest.mark.parametrize("radix", [0, 8])
def test_sort_stress_full_size_sort_properties(radix):
    """radix 0 sorts the 33-bit key as 9 + 8 + 8 + 8 bits, radix 8 as 7 + 7 + 7 + 6 + 6."""
    cfg = scenes.workload("sort-stress-6M-2160p")
    s = raster.RenderSettings(width=cfg.width, height=cfg.height, radix_bits=radix)
    rast = raster.TileRasterizer()
    targets = raster.RenderTargets(cfg.width, cfg.height)
    st = rast.render(cfg.camera, [raster.SplatInstance(cfg.scene.to_device())], s, targets)
    n, visible, pairs, tiles = (int(v) for v in rast.debug(_abi.DEBUG_COUNTS))
    print(f"sort stress: {n} splats, {visible} drawn, {pairs} pairs, {tiles} tiles, key {st.key_bits} bits, "
          f"{st.sort_passes} passes", flush=True)
    assert pairs == st.pairs and pairs > 2 ** 28
    touched = rast.debug(_abi.DEBUG_TOUCHED)
    rects = rast.debug(_abi.DEBUG_RECTS).reshape(-1, 2)
    ranges = rast.debug(_abi.DEBUG_RANGES).reshape(-1, 2).astype(np.int64)
    keys = rast.debug(_abi.DEBUG_KEYS)
    values = rast.debug(_abi.DEBUG_VALUES)
    print("  read back", flush=True)
    assert keys.size == pairs and values.size == pairs
    tiles_x = (cfg.width + 15) // 16
    x0, y0 = rects[:, 0] & 0xFFFF, rects[:, 0] >> 16
    x1, y1 = rects[:, 1] & 0xFFFF, rects[:, 1] >> 16
    prev = None
    run_start = np.zeros(tiles, np.int64)
    per_slot = np.zeros(n, np.int64)
    counts = np.zeros(tiles, np.int64)
    for c0 in range(0, pairs, CHUNK):
        k = keys[c0:c0 + CHUNK]
        v = values[c0:c0 + CHUNK].astype(np.int64)
        tile = (k >> np.uint64(32)).astype(np.int64)
        dk = ((k & np.uint64(0x7FFFFFFF)) >> np.uint64(7)).astype(np.int64)
        comp = tile << 24 | dk
        if prev is not None:
            comp_p = np.concatenate([[prev[0]], comp])
            v_p = np.concatenate([[prev[1]], v])
        else:
            comp_p, v_p = comp, v
        step = np.diff(comp_p)
        assert (step >= 0).all(), "pairs out of (tile, depth key) order"
        same = step == 0
        assert (np.diff(v_p)[same] > 0).all(), "equal keys out of slot order"
        prev = (int(comp[-1]), int(v[-1]))
        tx, ty = tile % tiles_x, tile // tiles_x
        assert ((tx >= x0[v]) & (tx <= x1[v]) & (ty >= y0[v]) & (ty <= y1[v])).all(), "a pair outside its rectangle"
        per_slot += np.bincount(v, minlength=n)[:n]
        counts += np.bincount(tile, minlength=tiles)[:tiles]
        print(f"  pairs [{c0}, {c0 + k.size}) ordered", flush=True)
    # Each slot as often as the tiles it touches; the runs cover every pair.
    assert np.array_equal(per_slot, touched.astype(np.int64))
    assert int(counts.sum()) == pairs
    run_start[1:] = np.cumsum(counts)[:-1]
    has = counts > 0
    assert np.array_equal(ranges[has, 0], run_start[has])
    assert np.array_equal(ranges[has, 1], run_start[has] + counts[has])
    assert (ranges[~has] == 0).all()
