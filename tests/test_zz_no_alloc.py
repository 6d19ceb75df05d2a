"""fg_radix_sort_pairs_noalloc: the sort on device buffers with its work in the library's
fixed device array, against fg_radix_sort_pairs (the same sort, allocating its own work) and
numpy's stable argsort, bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_PAIRS = 1 << 20


def hip():
    """The HIP runtime the library itself links (a second runtime in one process, such as
    torch's own, does not see the device once this one holds it)."""
    import ctypes as C
    import d3fgpu
    d3fgpu.lib()
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return h


class DeviceU32:
    """n u32 in device memory, for the test's own inputs and outputs."""
    H2D, D2H = 1, 2

    def __init__(self, n, host=None):
        import ctypes as C
        self.h, self.n, self.p = hip(), n, C.c_void_p()
        assert self.h.hipMalloc(C.byref(self.p), 4 * max(n, 1)) == 0
        if host is not None:
            a = np.ascontiguousarray(host, dtype=np.uint32)
            assert self.h.hipMemcpy(self.p, a.ctypes.data, 4 * n, self.H2D) == 0

    def read(self):
        out = np.zeros(self.n, np.uint32)
        assert self.h.hipMemcpy(out.ctypes.data, self.p, 4 * self.n, self.D2H) == 0
        return out

    def free(self):
        self.h.hipFree(self.p)


def run_noalloc(e, keys, values, first_bit=0, bits=32, in_place=False):
    n = len(keys)
    k, v = DeviceU32(n, keys), DeviceU32(n, values)
    ko, vo = (k, v) if in_place else (DeviceU32(n), DeviceU32(n))
    try:
        e.radix_sort_pairs_noalloc(n, k.p.value, v.p.value, ko.p.value, vo.p.value, first_bit, bits)
        return ko.read(), vo.read(), k.read(), v.read()
    finally:
        for b in {id(x): x for x in (k, v, ko, vo)}.values():
            b.free()


def check_against_allocating_sort(e, keys, first_bit=0, bits=32):
    values = np.arange(len(keys), dtype=np.uint32)
    ko, vo, k_in, v_in = run_noalloc(e, keys, values, first_bit, bits)
    rk, rv = e.radix_sort_pairs(keys, values, first_bit, bits)
    np.testing.assert_array_equal(ko, rk)
    np.testing.assert_array_equal(vo, rv)
    mask = (1 << (first_bit + bits)) - 1 if first_bit + bits < 32 else 0xffffffff
    np.testing.assert_array_equal(vo, np.argsort((keys.astype(np.uint64) & mask) >> first_bit, kind="stable"))
    np.testing.assert_array_equal(k_in, keys)  # the inputs are left as they were
    np.testing.assert_array_equal(v_in, values)


def test_two_calls_in_a_row_second_smaller(engine_factory):
    e = engine_factory()
    rng = np.random.default_rng(11)
    check_against_allocating_sort(e, rng.integers(0, 1 << 32, 300_000, dtype=np.uint64).astype(np.uint32))
    # smaller, many ties: a stale look-back word or histogram from the first call would show here
    check_against_allocating_sort(e, rng.integers(0, 37, 5_000, dtype=np.uint32))


def test_digit_range_and_tile_edges(engine_factory):
    e = engine_factory()
    rng = np.random.default_rng(12)
    for n in (1, 2, 3071, 3072, 3073):
        check_against_allocating_sort(e, rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32))
    check_against_allocating_sort(e, rng.integers(0, 1 << 32, 5_000, dtype=np.uint64).astype(np.uint32), 8, 16)


def test_largest_size_at_both_radix_widths(engine_factory, monkeypatch):
    rng = np.random.default_rng(13)
    keys = rng.integers(0, 1 << 32, MAX_PAIRS, dtype=np.uint64).astype(np.uint32)
    check_against_allocating_sort(engine_factory(), keys)
    monkeypatch.setenv("D3FGPU_RADIX_BITS", "11")  # read by fg_create
    check_against_allocating_sort(engine_factory(), keys[: MAX_PAIRS - 5])


def test_in_place(engine_factory):
    e = engine_factory()
    keys = np.random.default_rng(14).integers(0, 1000, 40_000, dtype=np.uint32)
    values = np.arange(len(keys), dtype=np.uint32)
    ko, vo, _, _ = run_noalloc(e, keys, values, in_place=True)
    ref = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(vo, ref)
    np.testing.assert_array_equal(ko, keys[ref])


def test_rejects_more_than_its_scratch_holds(engine_factory):
    import d3fgpu
    e = engine_factory()
    with pytest.raises(d3fgpu.FgError, match="1048577 pairs exceed the 1048576"):
        # rejected before any pointer is used
        e.radix_sort_pairs_noalloc(MAX_PAIRS + 1, 16, 16, 16, 16)
    # the context stays usable
    check_against_allocating_sort(e, np.arange(1000, 0, -1, dtype=np.uint32))
