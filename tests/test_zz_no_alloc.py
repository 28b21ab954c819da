"""tt_mst_deaths_dev_noalloc, the entry point that allocates no device memory,
against tt_mst_deaths_dev (its allocating counterpart) through the C ABI: the
sorted H0 deaths must be bitwise equal, on two calls in a row (the second on
a smaller cloud, so stale scratch contents from the first would show), and n
beyond the static scratch region must fail on the host with TT_ELENGTH."""
import ctypes as C

import numpy as np
import pytest

from turbotops_amd import capi

pytestmark = pytest.mark.gpu


def _stream(torch):
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _deaths(torch, lib, name, x, kind, p):
    n, dim = x.shape
    xd = torch.from_numpy(x).to("cuda:0")
    out = torch.full((max(n - 1, 1),), -1.0, dtype=torch.float32, device="cuda:0")
    if name == "tt_mst_deaths_dev":
        lib.call(name, C.c_void_p(xd.data_ptr()), n, dim, kind, p, C.c_void_p(out.data_ptr()), None, None, None,
                 _stream(torch))
    else:
        lib.call(name, C.c_void_p(xd.data_ptr()), n, dim, kind, p, C.c_void_p(out.data_ptr()), _stream(torch))
    torch.cuda.synchronize()
    return out.cpu().numpy()[: max(n - 1, 0)]


@pytest.mark.parametrize("kind,p", [(capi.METRIC_L2, 2.0), (capi.METRIC_L1, 1.0), (capi.METRIC_LINF, np.inf)])
def test_noalloc_matches_allocating_entry_twice(gpu, kind, p):
    import torch
    lib = capi.lib()
    rng = np.random.default_rng(11 + kind)
    for n, dim in ((5000, 3), (1200, 8)):   # second call smaller: stale scratch would show
        x = rng.random((n, dim), dtype=np.float32)
        print(f"no-alloc MST: kind {kind} n {n} dim {dim}", flush=True)
        want = _deaths(torch, lib, "tt_mst_deaths_dev", x, kind, p)
        got = _deaths(torch, lib, "tt_mst_deaths_dev_noalloc", x, kind, p)
        assert got.shape == (n - 1,)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
        assert np.all(np.diff(got) >= 0)


def test_noalloc_rejects_n_beyond_scratch(gpu):
    import torch
    lib = capi.lib()
    n = (1 << 20) + 1
    x = torch.zeros((n, 3), dtype=torch.float32, device="cuda:0")
    out = torch.zeros(n, dtype=torch.float32, device="cuda:0")
    rc = lib.dll.tt_mst_deaths_dev_noalloc(C.c_void_p(x.data_ptr()), n, 3, capi.METRIC_L2, 2.0,
                                           C.c_void_p(out.data_ptr()), _stream(torch))
    assert rc == capi.TT_ELENGTH
    assert "exceeds 1048576 points" in lib.last_error()


def test_noalloc_at_the_size_limit(gpu):
    """n = 2^20, the largest cloud the region holds (the H0 configuration's
    size class): equal to the allocating entry point."""
    import torch
    lib = capi.lib()
    rng = np.random.default_rng(4)
    x = rng.random((1 << 20, 3), dtype=np.float32)
    print("no-alloc MST: n 2^20 dim 3", flush=True)
    want = _deaths(torch, lib, "tt_mst_deaths_dev", x, capi.METRIC_L2, 2.0)
    got = _deaths(torch, lib, "tt_mst_deaths_dev_noalloc", x, capi.METRIC_L2, 2.0)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_noalloc_rejects_temporaries_beyond_scratch(gpu):
    """n = 2^20 in 20 dimensions needs more temporaries (the sorted copy of
    the points grows with dim) than the static region holds: TT_ELENGTH
    before any kernel runs."""
    import torch
    lib = capi.lib()
    n, dim = 1 << 20, 20
    x = torch.zeros((n, dim), dtype=torch.float32, device="cuda:0")
    out = torch.zeros(n, dtype=torch.float32, device="cuda:0")
    rc = lib.dll.tt_mst_deaths_dev_noalloc(C.c_void_p(x.data_ptr()), n, dim, capi.METRIC_L2, 2.0,
                                           C.c_void_p(out.data_ptr()), _stream(torch))
    assert rc == capi.TT_ELENGTH
    assert "scratch region too small" in lib.last_error()
