import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: runs HIP kernels on an MI355X (gfx950)")


@pytest.fixture(scope="session")
def restatement():
    from oracle import Restatement
    return Restatement()


@pytest.fixture(scope="session")
def reference():
    from oracle import Reference, reference_available
    if not reference_available():
        pytest.skip("oracle/_ref not built (reference sources absent at build time)")
    return Reference(simd=False)


@pytest.fixture(scope="session")
def golden():
    def load(name):
        return np.load(os.path.join(GOLDEN, name), allow_pickle=False)
    return load


@pytest.fixture(scope="session")
def gpu():
    """The HIP library on a visible GPU; skips only where no GPU exists."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    from turbotops_amd import capi
    return capi.lib()   # raises ImportError if libttgpu.so was not built


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def ulp_diff(a, b):
    """|a - b| in units in the last place (float32, finite, non-negative)."""
    ia = bits(a).astype(np.int64)
    ib = bits(b).astype(np.int64)
    return np.abs(ia - ib)
