"""The boundary (RedZeD) engine's edge list on the GPU (tt_build_edges,
tt_build_edges_csr) against the reference's own build_edges
(include/turbotops/redzed.hpp:544-613, compiled into oracle/_ref) and the
order's closed form (numpy lexsort of (diameter, combinatorial index))."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _comb_sorted(mat_full, thr):
    n = mat_full.shape[0]
    j, i = np.tril_indices(n, -1)           # j > i: rows j, columns i < j
    d = mat_full[j, i]
    keep = d <= thr
    i, j, d = i[keep], j[keep], d[keep]
    comb = j.astype(np.int64) * (j - 1) // 2 + i
    o = np.lexsort((comb, d))
    return np.stack([i[o], j[o]], 1).astype(np.int32), d[o]


def _grid(n_side, dim):
    g = np.stack(np.meshgrid(*[np.arange(n_side)] * dim, indexing="ij"), -1).reshape(-1, dim)
    return g.astype(np.float32)


CASES = [("gauss", 400, 3, "l2", math.inf), ("gauss", 333, 5, "l1", math.inf), ("grid", 7, 3, "linf", math.inf),
         ("grid", 6, 3, "l2", 2.0), ("uniform", 256, 8, "l2", 0.9), ("gauss", 2, 3, "l2", math.inf)]


@pytest.mark.parametrize("kind,n,dim,metric,thr", CASES)
def test_build_edges_vs_reference(gpu, reference, kind, n, dim, metric, thr):
    from turbotops_amd import build_edges
    rng = np.random.default_rng(n + dim)
    x = {"gauss": lambda: rng.normal(size=(n, dim)), "uniform": lambda: rng.uniform(size=(n, dim)),
         "grid": lambda: _grid(n, dim)}[kind]().astype(np.float32)
    print(f"case build_edges {kind} n={x.shape[0]} dim={dim} {metric} thr={thr}", flush=True)
    v, d, t, mat = build_edges(x, metric, threshold=thr, with_matrix=True)
    rv, rd, rt = reference.build_edges(x, metric, thr)
    assert np.float32(t) == np.float32(rt)
    assert np.array_equal(v, rv)
    assert np.array_equal(d.view(np.uint32), rd.view(np.uint32))
    cv, cd = _comb_sorted(mat, t)
    assert np.array_equal(v, cv) and np.array_equal(d, cd)


def test_build_edges_lp_and_lower_layout(gpu):
    from turbotops_amd import build_edges, make_dense
    x = np.random.default_rng(5).uniform(size=(300, 16)).astype(np.float32)
    v, d, t, low = build_edges(x, "lp:1.5", layout="lower", with_matrix=True)
    _, full, radius = make_dense(x, "lp:1.5")
    assert np.float32(t) == np.float32(radius)
    assert np.array_equal(low, full[np.tril_indices(300, -1)])
    cv, cd = _comb_sorted(full, t)
    assert np.array_equal(v, cv) and np.array_equal(d, cd)


def _tie_graph(seed, n, m):
    rng = np.random.default_rng(seed)
    rows = [dict() for _ in range(n)]
    for a, b, w in zip(rng.integers(0, n, m), rng.integers(0, n, m), rng.integers(1, 5, m)):
        if a != b:
            rows[a][b] = rows[b][a] = np.float32(w)
    rp = np.zeros(n + 1, np.uint64)
    col, ww = [], []
    for v in range(n):
        ks = sorted(rows[v])
        col += ks
        ww += [rows[v][k] for k in ks]
        rp[v + 1] = rp[v] + len(ks)
    return rp, np.array(col, np.uint32), np.array(ww, np.float32)


@pytest.mark.parametrize("case", ["ties", "knn", "radius"])
def test_build_edges_csr_vs_reference(gpu, reference, case):
    from turbotops_amd import build_edges_csr, knn_graph_exact, radius_graph
    if case == "ties":
        n = 80
        rp, col, w = _tie_graph(3, n, 900)
        thr = 3.0
    else:
        x = np.random.default_rng(9).uniform(size=(2000, 3)).astype(np.float32)
        g = knn_graph_exact(x, "l2", 10) if case == "knn" else radius_graph(x, "l2", 0.1)
        n, rp, col, w, thr = g.n, g.row_ptr, g.col_idx, g.weights, math.inf
    print(f"case build_edges_csr {case} n={n} nnz={int(rp[-1])}", flush=True)
    v, d = build_edges_csr((n, rp, col, w), threshold=thr)
    rv, rd = reference.build_edges_csr(n, rp, col, w, thr)
    assert np.array_equal(v, rv) and np.array_equal(d.view(np.uint32), rd.view(np.uint32))
    assert v.shape[0] > 0 and (v[:, 0] < v[:, 1]).all()


@pytest.mark.parametrize("kind,n,dim,metric", [("gauss", 700, 3, "l2"), ("grid", 7, 3, "linf"), ("grid", 8, 3, "l1"),
                                                ("uniform", 300, 8, "l2"), ("gauss", 2, 3, "l2")])
def test_witness_flags_vs_definition(gpu, kind, n, dim, metric):
    """tt_build_edges_witness: the same list as tt_build_edges, and every
    edge's ready-witness flag equal to its definition (numpy, every w). The
    grids put hundreds of equal edges in one diameter; n = 700 reaches the
    block-filtered second phase of the search."""
    from test_boundary_engine_cpu import witness_flags
    from turbotops_amd import build_edges
    rng = np.random.default_rng(3 * n + dim)
    x = {"gauss": lambda: rng.normal(size=(n, dim)), "uniform": lambda: rng.uniform(size=(n, dim)),
         "grid": lambda: _grid(n, dim)}[kind]().astype(np.float32)
    print(f"case witness {kind} n={x.shape[0]} dim={dim} {metric}", flush=True)
    v, d, t, mat = build_edges(x, metric, with_matrix=True)
    vw, dw, flags, tw = build_edges(x, metric, witness=True)
    assert np.array_equal(v, vw) and np.array_equal(d.view(np.uint32), dw.view(np.uint32)) and t == tw
    want = witness_flags(mat, v.astype(np.uint32), d)
    assert np.array_equal(flags, want)


def test_witness_flags_sampled_large(gpu):
    """n = 4200 (more than one pass over the block minima): 30000 sampled
    edges against the definition."""
    from test_boundary_engine_cpu import witness_flags
    from turbotops_amd import build_edges
    x = np.random.default_rng(17).normal(size=(4200, 3)).astype(np.float32)
    print("case witness sampled n=4200", flush=True)
    v, d, flags, t, mat = build_edges(x, "l2", witness=True, with_matrix=True)
    pick = np.random.default_rng(1).choice(len(d), 30000, replace=False)
    want = witness_flags(mat, v[pick].astype(np.uint32), d[pick])
    assert np.array_equal(flags[pick], want)
    assert flags.mean() > 0.9
