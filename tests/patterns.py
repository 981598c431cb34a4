"""The reference's GPU test inputs, made on the host: sort_generate's five key
patterns (shaders/lrt/test/sort_generate.slang:24-49, seeds as
tests/gpu/test_algo.cpp:63) and counts_generate's scan input
(shaders/lrt/test/counts_generate.slang:23-32)."""
import numpy as np

M32 = np.uint64(0xFFFFFFFF)


def pcg(v: np.ndarray) -> np.ndarray:
    v = v.astype(np.uint64) & M32
    state = (v * np.uint64(747796405) + np.uint64(2891336453)) & M32
    word = (((state >> ((state >> np.uint64(28)) + np.uint64(4))) ^ state) * np.uint64(277803737)) & M32
    return ((word >> np.uint64(22)) ^ word) & M32


def sort_generate(count: int, mode: int, key_bits: int):
    """(keys u64, values u32): keys = hi << 32 | lo, masked to key_bits; values = index."""
    seed = np.uint64((0x9E3779B9 ^ count ^ (mode << 8)) & 0xFFFFFFFF)
    i = np.arange(count, dtype=np.uint64)
    if mode == 0:
        lo = pcg(i ^ seed)
        hi = pcg((lo + seed) & M32)
    elif mode == 1:
        lo = np.full(count, 0x5A5A5A5A, np.uint64)
        hi = np.full(count, 0x0000A5A5, np.uint64)
    elif mode == 2:
        lo = i.copy()
        hi = np.zeros(count, np.uint64)
    elif mode == 3:
        lo = np.uint64(count - 1) - i
        hi = np.full(count, 7, np.uint64)
    else:
        lo = pcg(i ^ seed) % np.uint64(13)
        hi = pcg((i + seed) & M32) % np.uint64(3)
    mask_lo = 0xFFFFFFFF if key_bits >= 32 else (1 << key_bits) - 1
    mask_hi = 0 if key_bits <= 32 else (0xFFFFFFFF if key_bits >= 64 else (1 << (key_bits - 32)) - 1)
    keys = ((hi & np.uint64(mask_hi)) << np.uint64(32)) | (lo & np.uint64(mask_lo))
    return keys.astype(np.uint64), np.arange(count, dtype=np.uint32)


def counts_generate(count: int, seed: int = 1234) -> np.ndarray:
    h = pcg(np.arange(count, dtype=np.uint64) ^ np.uint64(seed))
    return np.where((h & np.uint64(7)) == 0, 0, h % np.uint64(37)).astype(np.uint32)


def stable_sorted(keys: np.ndarray, values: np.ndarray, key_bits: int):
    mask = np.uint64((1 << key_bits) - 1) if key_bits < 64 else np.uint64(0xFFFFFFFFFFFFFFFF)
    order = np.argsort(keys & mask, kind="stable")
    return keys[order], values[order]


# (count, keyBits) of tests/gpu/test_algo.cpp:357-358
SORT_CASES = [(1, 32), (2, 32), (4097, 13), (65537, 32), (300007, 24), (250001, 64), (100000, 40)]
SCAN_SIZES = [1, 7, 4095, 4096, 4097, 100003, 1 << 22]
