"""The radix sort (turbotops_amd/csrc/sort.hip) checked directly against a
stable sort on the host: every edge order, graph and gradient in the library
goes through it, so its stability and its handling of partial tiles, uniform
digits and the device-side count are pinned here on their own.

The downsweep ranks a wave's keys from a per-wave LDS slot per digit, or takes
the whole wave when its digits are uniform; the cases below reach both (random
digits, one digit, two digits spread over the lanes)."""
import ctypes as C
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libsort_device.so")
_u64p = C.POINTER(C.c_uint64)
_u32p = C.POINTER(C.c_uint32)


@pytest.fixture(scope="module")
def sort(gpu):
    L = C.CDLL(_LIB)
    L.sd_sort.argtypes = [_u64p, _u32p, C.c_uint64, C.c_int, C.c_int, C.c_int, C.c_uint64, _u64p, _u32p]
    L.sd_sort.restype = C.c_int

    def run(keys, lo, hi, vals=None, counted=False, capacity=0):
        keys = np.ascontiguousarray(keys, np.uint64)
        n = keys.size
        out_k = np.empty(n, np.uint64)
        out_v = np.empty(n, np.uint32) if vals is not None else None
        v = np.ascontiguousarray(vals, np.uint32) if vals is not None else None
        rc = L.sd_sort(keys.ctypes.data_as(_u64p), v.ctypes.data_as(_u32p) if v is not None else None, n, lo, hi,
                       1 if counted else 0, max(capacity, n), out_k.ctypes.data_as(_u64p),
                       out_v.ctypes.data_as(_u32p) if out_v is not None else None)
        assert rc == 0
        return out_k, out_v
    return run


def _stable(keys, lo, hi):
    field = (keys >> np.uint64(lo)) & np.uint64((1 << (hi - lo)) - 1) if hi - lo < 64 else keys
    return np.argsort(field, kind="stable")


def _cases():
    rng = np.random.default_rng(11)
    r64 = lambda n: rng.integers(0, 2 ** 63, size=n, dtype=np.uint64) * np.uint64(2) + \
        rng.integers(0, 2, size=n, dtype=np.uint64)
    yield "random-64-bit", r64(300001), 0, 64
    yield "random-field", r64(100000), 13, 47
    yield "one-digit", np.full(50000, 0x1234_5678_9abc, np.uint64) ^ rng.integers(0, 1 << 12, size=50000, dtype=np.uint64), 12, 40
    yield "two-digits", (rng.integers(0, 2, size=70000, dtype=np.uint64) << np.uint64(20)) | \
        rng.integers(0, 1 << 20, size=70000, dtype=np.uint64), 20, 28
    yield "few-values", rng.integers(0, 3, size=40000, dtype=np.uint64) * np.uint64(0x0101_0101), 0, 32
    yield "tile-plus-one", r64(4097), 0, 40
    yield "under-one-wave", r64(37), 0, 24
    yield "sorted", np.arange(20000, dtype=np.uint64) * np.uint64(7919), 0, 40
    yield "reversed", np.arange(20000, 0, -1, dtype=np.uint64), 0, 16


@pytest.mark.parametrize("name,keys,lo,hi", list(_cases()), ids=[c[0] for c in _cases()])
def test_keys_and_payload_match_stable_sort(sort, name, keys, lo, hi):
    print(f"case {name} n={keys.size} bits [{lo}, {hi})", flush=True)
    vals = np.arange(keys.size, dtype=np.uint32)
    order = _stable(keys, lo, hi)
    got_k, got_v = sort(keys, lo, hi, vals=vals)
    assert np.array_equal(got_k, keys[order])
    assert np.array_equal(got_v, vals[order])          # stability: equal fields keep input order
    got_k2, _ = sort(keys, lo, hi)
    assert np.array_equal(got_k2, keys[order])


@pytest.mark.parametrize("n,capacity,lo,hi", [(100000, 160000, 0, 40), (4096 * 3 + 5, 4096 * 3 + 5, 8, 32),
                                              (777, 1 << 20, 0, 24), (50000, 50000, 3, 11)])
def test_counted_sort_reads_the_count_on_the_device(sort, n, capacity, lo, hi):
    """radix_sort_u64_counted: the capacity sizes the grid, the count comes
    from device memory; an odd pass count hands back the other buffer."""
    print(f"case counted n={n} capacity={capacity} bits [{lo}, {hi})", flush=True)
    keys = np.random.default_rng(n).integers(0, 2 ** 62, size=n, dtype=np.uint64)
    got, _ = sort(keys, lo, hi, counted=True, capacity=capacity)
    assert np.array_equal(got, keys[_stable(keys, lo, hi)])


def test_planted_unstable_order_is_caught():
    """The stability check can fail: a sort by the field that reverses the
    order of equal fields differs from the stable order."""
    keys = np.array([5, 1, 5, 1, 5], np.uint64) << np.uint64(8) | np.arange(5, dtype=np.uint64)
    order = _stable(keys, 8, 16)
    bad = np.lexsort((-np.arange(5), (keys >> np.uint64(8))))
    assert not np.array_equal(order, bad)
