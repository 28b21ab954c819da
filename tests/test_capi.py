"""The C ABI library loads, exports what include/ttgpu.h declares, and its
host-only entry points behave like the reference (no GPU needed)."""
import ctypes as C

import pytest

from turbotops_amd import capi


def test_exports_every_declared_symbol():
    lib = capi.lib()
    declared = capi.declared_symbols()
    assert len(declared) >= 7
    missing = [s for s in declared if not lib.has(s)]
    assert not missing, missing


def test_version_matches_reference():
    assert capi.lib().dll.tt_version().decode() == "0.1.0"


@pytest.mark.parametrize("s,kind,p", [("l2", 1, 2.0), ("euclidean", 1, 2.0), ("l1", 0, 1.0),
                                      ("manhattan", 0, 1.0), ("linf", 2, float("inf")),
                                      ("chebyshev", 2, float("inf")), ("lp:inf", 2, float("inf")),
                                      ("lp:1.5", 3, 1.5), ("lp:2", 1, 2.0), ("lp:1", 0, 1.0)])
def test_metric_parse(s, kind, p):
    from turbotops_amd import parse_metric
    assert parse_metric(s) == (kind, p)


@pytest.mark.parametrize("s,msg", [("lp:0.5", "metric: p must be in [1, inf]"),
                                   ("cosine", "metric: unknown 'cosine'"),
                                   ("lp:", "metric: unknown"), ("lp:nan", "p must be in")])
def test_metric_errors(s, msg):
    from turbotops_amd import parse_metric
    with pytest.raises(ValueError, match=msg.replace("[", r"\[").replace("]", r"\]")):
        parse_metric(s)


def test_dense_layout_gate():
    from turbotops_amd import dense_layout
    # rips.hpp:395-414; tests/test_cli.sh:143-148 (n=30: full 3600 B, lower 1740 B)
    assert dense_layout(30, 3600) == "full"
    assert dense_layout(30, 3000) == "lower"
    assert dense_layout(30, 1740) == "lower"
    with pytest.raises(ValueError, match="exceeds the memory budget"):
        dense_layout(30, 500)
    assert dense_layout(0, 0) == "full"


def test_errors_are_thread_local_and_status_mapped():
    lib = capi.lib()
    k, p = C.c_int(), C.c_float()
    rc = lib.dll.tt_metric_parse(b"wat", C.byref(k), C.byref(p))
    assert rc == capi.TT_EINVAL
    assert "unknown 'wat'" in lib.last_error()


def test_backward_host_form_rejects_out_of_range_endpoint():
    """tt_diagram_backward checks every endpoint on the host before any
    device work: an endpoint >= n is TT_EINVAL."""
    import numpy as np
    lib = capi.lib()
    x = np.zeros((4, 2), dtype=np.float32)
    edges = np.array([0, 1, 2, 4], dtype=np.uint32)   # second edge ends at 4 == n
    g = np.ones(2, dtype=np.float32)
    dx = np.zeros_like(x)
    rc = lib.dll.tt_diagram_backward(x.ctypes.data, 4, 2, capi.METRIC_L2, 2.0, edges.ctypes.data, g.ctypes.data, 2,
                                     dx.ctypes.data)
    assert rc == capi.TT_EINVAL
    assert lib.last_error() == "tt_diagram_backward: edge endpoint out of range"
