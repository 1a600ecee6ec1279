"""gsplat_sort_values_by_keys_noalloc through the C ABI: the allocation-free
variant must give the result of gsplat_sort_values_by_keys bit for bit, on two
calls in a row (the second shorter, so stale contents of the shared device
array cannot pass unnoticed), and refuse lengths beyond its 2^20 limit."""

import sys

import numpy as np
import pytest

from helpers import lcg_keys_fast

pytestmark = pytest.mark.gpu
LIMIT = 1 << 20


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def noalloc_sort(abi, keys, values):
    f = abi.lib().gsplat_sort_values_by_keys_noalloc
    k = np.ascontiguousarray(keys, dtype=np.uint32)
    v = np.ascontiguousarray(values, dtype=np.uint32).copy()
    rc = f(k.ctypes.data if k.size else None, v.ctypes.data if v.size else None, k.size)
    return rc, v


def test_noalloc_sort_equals_allocating_sort_twice(gpu, oracle):
    rng = np.random.default_rng(77)
    for case, n in (("first", LIMIT), ("second, shorter", 300_001)):
        keys = lcg_keys_fast(n, 1000 + n)
        keys[rng.integers(0, n, n // 8)] = 7  # ties: stability shows
        vals = rng.permutation(n).astype(np.uint32)
        say(f"[noalloc sort] {case} call n={n}")
        rc, got = noalloc_sort(gpu, keys, vals)
        assert rc == 0, gpu.last_error()
        want = gpu.sort_values_by_keys(keys, vals)
        assert np.array_equal(got, want), f"{case} call differs from gsplat_sort_values_by_keys"
        if n <= 300_001:
            assert np.array_equal(got, oracle.sort_values_by_keys(keys, vals))


def test_noalloc_sort_small_and_refused_lengths(gpu):
    for n in (0, 1, 2, 257, 4099):
        keys = lcg_keys_fast(n, 5 + n)
        vals = np.arange(n, dtype=np.uint32)
        say(f"[noalloc sort] n={n}")
        rc, got = noalloc_sort(gpu, keys, vals)
        assert rc == 0 and np.array_equal(got, gpu.sort_values_by_keys(keys, vals)), n
    rc, _ = noalloc_sort(gpu, np.zeros(LIMIT + 1, np.uint32), np.zeros(LIMIT + 1, np.uint32))
    assert rc == gpu.UNSUPPORTED and "no-allocation limit" in gpu.last_error()
