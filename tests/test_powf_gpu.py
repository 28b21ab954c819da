"""The device build of the Lp kernels' powers (turbotops_amd/csrc/powf.hpp)
over EVERY float of the ranges the kernels use, on the GPU: terms x^p for x
in (0, 1] and roots s^(1/p) for s in [1, 16], against the ROCm device
library's double-precision power rounded once to float. At most one ulp
apart, and equal in all but a few per million (near-midpoint cases of the
2^-43 relative error bound). Together with the host sweep against glibc's
powf (oracle/powf_sweep.cpp), this keeps every Lp term within one ulp of the
reference's powf."""
import ctypes as C
import os

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sweep(gpu):
    path = os.path.join(ROOT, "tests", "libpowf_device.so")
    if not os.path.exists(path):
        pytest.skip("tests/libpowf_device.so not built (make -C tests)")
    L = C.CDLL(path)
    L.pd_sweep.argtypes = [C.c_uint32, C.c_uint32, C.c_float, C.c_uint, C.POINTER(C.c_ulonglong),
                           C.POINTER(C.c_uint)]

    def run(lo, hi, p, half=0):
        mm, mu = C.c_ulonglong(), C.c_uint()
        assert L.pd_sweep(lo, hi, p, half, C.byref(mm), C.byref(mu)) == 0
        return mm.value, mu.value
    return run


def _bits(v):
    return int(np.array([v], np.float32).view(np.uint32)[0])


@pytest.mark.parametrize("p", [1.25, 2.7, 7.5, 50.0])
def test_pow_terms_every_float(sweep, p):
    mism, mu = sweep(1, _bits(1.0), p)
    print(f"pow_f32 terms p={p}: {mism} of {_bits(1.0)} differ, max {mu} ulp", flush=True)
    assert mu <= 1 and mism < 4e-6 * _bits(1.0)


@pytest.mark.parametrize("p", [1.5, 1.25, 3.3, 50.0])
def test_pow_roots_every_float(sweep, p):
    q = float(np.float32(1.0) / np.float32(p))
    lo, hi = _bits(1.0), _bits(16.0)
    mism, mu = sweep(lo, hi, q)
    print(f"pow_f32 roots 1/p={q}: {mism} of {hi - lo + 1} differ, max {mu} ulp", flush=True)
    assert mu <= 1 and mism < 4e-6 * (hi - lo + 1)


@pytest.mark.parametrize("n", [1, 2, 3, 7])
def test_half_integral_terms_every_float(sweep, n):
    mism, mu = sweep(1, _bits(1.0), 0.0, n)
    print(f"pow_half_f32 n={n}: {mism} of {_bits(1.0)} differ, max {mu} ulp", flush=True)
    assert mu <= 1 and mism < 1e-6 * _bits(1.0)
