"""GPU parity of the dimension-0 edge pipeline (tt_dense_edges /
tt_dense_edges_dev): kept edges in filtration order and apparent flags,
bit-exact against the oracle (and so against Reducer::dim0,
rips.hpp:685-849)."""
import ctypes as C

import numpy as np
import pytest

from conftest import bits

pytestmark = pytest.mark.gpu


def _edges(x, metric, threshold=np.inf, apparent=True):
    from turbotops_amd import dense_edges
    return dense_edges(x, metric, threshold, apparent)


@pytest.mark.parametrize("metric", ["l2", "l1", "linf"])
@pytest.mark.parametrize("n,d,kind", [(2, 3, "gauss"), (3, 2, "gauss"), (40, 3, "gauss"), (65, 3, "grid"),
                                      (130, 2, "grid"), (200, 4, "gauss"), (257, 16, "unif")])
def test_dense_edges_match_oracle(gpu, restatement, metric, n, d, kind):
    rng = np.random.default_rng(n * 7 + d)
    x = {"gauss": lambda: rng.normal(size=(n, d)), "unif": lambda: rng.uniform(size=(n, d)),
         "grid": lambda: rng.integers(0, 4, size=(n, d))}[kind]().astype(np.float32)
    print(f"case metric={metric} n={n} d={d} {kind}", flush=True)
    mat = restatement.make_dense(x, metric, True)
    thr = restatement.enclosing_radius(mat, n)
    idx, dd = restatement.dense_edges(mat, thr)
    flags = restatement.apparent_edges(mat, idx, dd)
    gi, gd, gf, gthr = _edges(x, metric)
    assert bits(np.float32(gthr)) == bits(np.float32(thr))
    assert np.array_equal(gi, idx)
    assert np.array_equal(bits(gd), bits(dd))
    assert np.array_equal(gf, flags)


def test_finite_threshold_and_ties(gpu, restatement):
    rng = np.random.default_rng(9)
    x = rng.integers(0, 3, size=(150, 3)).astype(np.float32)
    mat = restatement.make_dense(x, "l1", True)
    for thr in [0.0, 1.0, 2.0, 100.0]:
        idx, dd = restatement.dense_edges(mat, thr)
        gi, gd, gf, _ = _edges(x, "l1", thr)
        assert np.array_equal(gi, idx) and np.array_equal(bits(gd), bits(dd))
        assert np.array_equal(gf, restatement.apparent_edges(mat, idx, dd))


def test_large_sorted_and_counts(gpu, restatement):
    """n = 5000: counts against the matrix, order against numpy's lexsort,
    and apparent flags on a sample against the oracle."""
    rng = np.random.default_rng(2)
    n = 5000
    x = rng.normal(size=(n, 3)).astype(np.float32)
    gi, gd, gf, thr = _edges(x, "l2")
    mat = restatement.make_dense(x, "l2", True)
    assert bits(np.float32(thr)) == bits(np.float32(restatement.enclosing_radius(mat, n)))
    ii, jj = np.tril_indices(n, -1)
    dv = mat[ii, jj]
    keep = dv <= thr
    assert gi.size == keep.sum()
    ref_idx = (ii.astype(np.uint64) * (ii.astype(np.uint64) - 1) // 2 + jj.astype(np.uint64))[keep]
    ref_d = dv[keep]
    order = np.lexsort((-ref_idx.astype(np.int64), ref_d))
    assert np.array_equal(gi, ref_idx[order])
    assert np.array_equal(bits(gd), bits(ref_d[order]))
    sample = rng.choice(gi.size, size=3000, replace=False)
    want = restatement.apparent_edges(mat, gi[sample], gd[sample])
    assert np.array_equal(gf[sample], want)


def test_device_entry_two_phase(gpu, restatement):
    import torch
    from turbotops_amd import capi, dist_matrix_dev
    rng = np.random.default_rng(4)
    x = rng.normal(size=(300, 3)).astype(np.float32)
    xt = torch.from_numpy(x).cuda()
    mat = dist_matrix_dev(xt, "l2", "full")
    m = C.c_uint64()
    lib = capi.lib()
    lib.call("tt_dense_edges_dev", C.c_void_p(mat.data_ptr()), 300, C.c_float(1.0), 1, 0, C.byref(m),
             None, None, None, None)
    idx = torch.empty(m.value, dtype=torch.int64, device="cuda")
    d = torch.empty(m.value, dtype=torch.float32, device="cuda")
    ap = torch.empty(m.value, dtype=torch.uint8, device="cuda")
    lib.call("tt_dense_edges_dev", C.c_void_p(mat.data_ptr()), 300, C.c_float(1.0), 1, m.value, C.byref(m),
             C.c_void_p(idx.data_ptr()), C.c_void_p(d.data_ptr()), C.c_void_p(ap.data_ptr()), None)
    torch.cuda.synchronize()
    full = restatement.make_dense(x, "l2", True)
    ri, rd = restatement.dense_edges(full, 1.0)
    assert np.array_equal(idx.cpu().numpy().astype(np.uint64), ri)
    assert np.array_equal(bits(d.cpu().numpy()), bits(rd))
    assert np.array_equal(ap.cpu().numpy(), restatement.apparent_edges(full, ri, rd))
    with pytest.raises(ValueError, match="capacity"):
        lib.call("tt_dense_edges_dev", C.c_void_p(mat.data_ptr()), 300, C.c_float(1.0), 1, 1, C.byref(m),
                 C.c_void_p(idx.data_ptr()), None, None, None)


@pytest.mark.parametrize("metric,n,thr", [("linf", 4200, 1.0), ("l2", 4161, 1.5), ("l1", 130, np.inf),
                                          ("l2", 700, np.inf)])
def test_block_filtered_apparent_all_edges(gpu, restatement, metric, n, thr):
    """The apparent scan reads only the 64-column blocks whose minima allow a
    cofacet; every flag must equal the plain walk of the oracle. Tie-heavy
    integer grids, n with a partial last block and n > 64 * 64 (more than one
    pass over the block minima)."""
    rng = np.random.default_rng(n)
    x = rng.integers(0, 9, size=(n, 3)).astype(np.float32)
    print(f"case block-filtered apparent metric={metric} n={n} thr={thr}", flush=True)
    gi, gd, gf, thr = _edges(x, metric, thr)
    mat = restatement.make_dense(x, metric, True)
    idx, dd = restatement.dense_edges(mat, thr)
    assert np.array_equal(gi, idx) and np.array_equal(bits(gd), bits(dd))
    want = restatement.apparent_edges(mat, idx, dd)
    assert np.array_equal(gf, want)
    assert 0 < gf.sum() < gf.size


def test_dense_edges_dev_wrapper(gpu, restatement):
    import torch
    from turbotops_amd import dense_edges_dev, dist_matrix_dev
    x = np.random.default_rng(8).normal(size=(700, 3)).astype(np.float32)
    mat = dist_matrix_dev(torch.from_numpy(x).cuda(), "l2", "full").view(700, 700)
    m, idx, d, ap = dense_edges_dev(mat, 1.5)
    assert dense_edges_dev(mat, 1.5, count_only=True) == m
    full = restatement.make_dense(x, "l2", True)
    ri, rd = restatement.dense_edges(full, 1.5)
    assert np.array_equal(idx.cpu().numpy().astype(np.uint64), ri)
    assert np.array_equal(bits(d.cpu().numpy()), bits(rd))
    assert np.array_equal(ap.cpu().numpy(), restatement.apparent_edges(full, ri, rd))
