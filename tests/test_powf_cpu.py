"""The Lp kernels' power (turbotops_amd/csrc/powf.hpp), host build of the
same source through tests/libhost_harness.so, against the C library's powf
(glibc, which the reference calls at include/turbotops/metric.hpp:204, :206)
on samples of the ranges the kernels use — terms x in (0, 1], roots s in
[1, 16] — and on closed forms: at most one ulp apart, and exact where x^p is
a float (powers of two, x = 1, x = 0). The exhaustive sweep over every float
in those ranges is oracle/powf_sweep.cpp."""
import ctypes as C
import ctypes.util
import os

import numpy as np
import pytest

from conftest import ROOT, ulp_diff

_f32p = C.POINTER(C.c_float)


@pytest.fixture(scope="module")
def hpow():
    path = os.path.join(ROOT, "tests", "libhost_harness.so")
    if not os.path.exists(path):
        pytest.skip("tests/libhost_harness.so not built (make -C tests)")
    L = C.CDLL(path)
    L.hh_pow_f32.argtypes = [_f32p, C.c_uint64, C.c_float, _f32p]
    L.hh_pow_half_f32.argtypes = [_f32p, C.c_uint64, C.c_uint32, _f32p]

    def run(x, p, half=None):
        x = np.ascontiguousarray(x, np.float32)
        out = np.empty_like(x)
        if half is None:
            L.hh_pow_f32(x.ctypes.data_as(_f32p), x.size, p, out.ctypes.data_as(_f32p))
        else:
            L.hh_pow_half_f32(x.ctypes.data_as(_f32p), x.size, half, out.ctypes.data_as(_f32p))
        return out
    return run


@pytest.fixture(scope="module")
def libm_powf():
    m = C.CDLL(ctypes.util.find_library("m"))
    m.powf.argtypes = [C.c_float, C.c_float]
    m.powf.restype = C.c_float
    return np.vectorize(lambda x, p: m.powf(float(x), float(p)), otypes=[np.float32])


@pytest.mark.parametrize("p", [1.5, 1.25, 2.5, 3.7, 7.5, 50.0, 1.1])
def test_terms_and_roots_within_one_ulp_of_libm(hpow, libm_powf, p):
    rng = np.random.default_rng(int(p * 100))
    # terms: uniform bit patterns over (0, 1] (every binade) plus uniform values
    bits = rng.integers(1, 0x3F800001, size=60000, dtype=np.int64).astype(np.uint32)
    x = np.concatenate([bits.view(np.float32), rng.uniform(0, 1, 40000).astype(np.float32)])
    x = x[x > 0]
    got, want = hpow(x, p), libm_powf(x, np.float32(p))
    assert ulp_diff(got, want).max() <= 1
    q = np.float32(1.0) / np.float32(p)   # the reference's 1.0f / p
    s = rng.uniform(1, 16, 100000).astype(np.float32)
    got, want = hpow(s, q), libm_powf(s, q)
    assert ulp_diff(got, want).max() <= 1


@pytest.mark.parametrize("n", [1, 2, 3, 7])
def test_half_integral_terms(hpow, libm_powf, n):
    """pow_half_f32 (x^n sqrt x in double, the Lp kernels' terms for p = n + 1/2)
    within one ulp of libm and equal to the double-precision power rounded once."""
    rng = np.random.default_rng(n)
    bits = rng.integers(1, 0x3F800001, size=80000, dtype=np.int64).astype(np.uint32)
    x = bits.view(np.float32)
    got = hpow(x, None, half=n)
    assert ulp_diff(got, libm_powf(x, np.float32(n + 0.5))).max() <= 1
    assert np.array_equal(got, np.power(x.astype(np.float64), n + 0.5).astype(np.float32))


def test_exact_cases(hpow):
    assert hpow(np.array([0.0], np.float32), 1.5)[0] == 0.0
    assert hpow(np.array([1.0], np.float32), 1.5)[0] == 1.0
    assert hpow(np.array([1.0], np.float32), np.float32(1) / np.float32(3))[0] == 1.0
    two = 2.0 ** -np.arange(0, 60, dtype=np.float64)
    got = hpow(two.astype(np.float32), 2.5)
    assert np.array_equal(got, (two ** 2.5).astype(np.float32))   # 2^(-2.5k): exact in double, one rounding
    got = hpow(np.array([4.0, 9.0, 16.0], np.float32), 0.5)
    assert np.array_equal(got, np.array([2.0, 3.0, 4.0], np.float32))
    assert hpow(np.array([1e-30], np.float32), 50.0)[0] == 0.0       # underflow to zero, as powf


def test_planted_errors_are_caught(hpow, libm_powf):
    """A power off by a relative 2^-20 (a dropped polynomial term would be
    far worse) fails the one-ulp check: the check can fail."""
    x = np.random.default_rng(3).uniform(0.01, 1, 20000).astype(np.float32)
    bad = (hpow(x, 1.5).astype(np.float64) * (1 + 2.0 ** -20)).astype(np.float32)
    assert ulp_diff(bad, libm_powf(x, np.float32(1.5))).max() > 1
