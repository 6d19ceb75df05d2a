"""The radix sort on its own: sorted on the requested digits and stable
(test/gpu-sort-test.js:9-81 is the model: same sizes around tile edges, same oracle)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def random_keys(n, spread, seed=7):
    """gpu-sort-test.js:142-146: s = (imul(s, 1664525) + 1013904223) >>> 0, key = s % spread."""
    s = seed
    out = np.zeros(n, np.uint64)
    for i in range(n):
        s = (s * 1664525 + 1013904223) & 0xffffffff
        out[i] = s % spread
    return out.astype(np.uint32)


def expected(keys, first_bit, bits):
    """Stable ascending on the masked digits, ties by input index (gpu-sort-test.js:130-136)."""
    mask = (1 << (first_bit + bits)) - 1 if first_bit + bits < 32 else 0xffffffff
    d = (keys.astype(np.uint64) & mask) >> first_bit
    return np.argsort(d, kind="stable")


@pytest.mark.parametrize("n", [1, 2, 3, 255, 256, 257, 1023, 1024, 1025, 3071, 3072, 3073,
                               3 * 3072 + 17, 100000])
def test_sorts_random_32bit_keys_stably(engine_factory, n):
    e = engine_factory()
    keys = random_keys(n, 4294967296)
    ko, vo = e.radix_sort_pairs(keys, np.arange(n, dtype=np.uint32), 0, 32)
    ref = expected(keys, 0, 32)
    np.testing.assert_array_equal(vo, ref)
    np.testing.assert_array_equal(ko, keys[ref])


def test_keeps_equal_keys_in_order(engine_factory):
    e = engine_factory()
    keys = random_keys(20000, 37)
    ko, vo = e.radix_sort_pairs(keys, np.arange(20000, dtype=np.uint32), 0, 32)
    np.testing.assert_array_equal(vo, expected(keys, 0, 32))


def test_sorts_on_requested_digits_only(engine_factory):
    e = engine_factory()
    keys = random_keys(5000, 4294967296)
    ko, vo = e.radix_sort_pairs(keys, np.arange(5000, dtype=np.uint32), 8, 16)
    np.testing.assert_array_equal(vo, expected(keys, 8, 16))


def test_large_sort_two_million(engine_factory):
    e = engine_factory()
    rng = np.random.default_rng(3)
    keys = rng.integers(0, 1 << 21, 2_000_000, dtype=np.uint32)
    ko, vo = e.radix_sort_pairs(keys, np.arange(len(keys), dtype=np.uint32), 0, 21)
    np.testing.assert_array_equal(vo, np.argsort(keys, kind="stable"))
