"""Dimension 0 of the reducer on sparse complexes (tt_sparse_edges):
Reducer::dim0's edge order and apparent test over SparseComplex::cofacets
(include/turbotops/rips.hpp:685-849, :428-471) on CSR graphs.

* Order: numpy's lexsort of (diameter asc, index desc) over the kept edges.
* Apparent flags: the definition, brute force (the first common neighbour x,
  descending, with max(w(a,x), w(b,x)) <= d is the zero-pivot cofacet; the
  edge is apparent iff it is that triangle's first facet of diameter d in
  ascending index order), on small graphs with many ties; and the oracle's
  restatement (tto_apparent_cx, pinned against the reference's
  zero_apparent_cofacet) on larger ones.
* End to end, the kNN and radius branches of persistence() already match the
  reference bar for bar (test_persistence_gpu.py); they run through this path.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _csr_from_edges(n, edges, w):
    rows = [dict() for _ in range(n)]
    for (a, b), x in zip(edges, w):
        if a == b:
            continue
        rows[a][b] = x
        rows[b][a] = x
    rp = np.zeros(n + 1, np.uint64)
    col, ww = [], []
    for v in range(n):
        ks = sorted(rows[v])
        col += ks
        ww += [rows[v][k] for k in ks]
        rp[v + 1] = rp[v] + len(ks)
    return rp, np.array(col, np.uint32), np.array(ww, np.float32)


def _random_graph(seed, n, m, ties):
    rng = np.random.default_rng(seed)
    e = rng.integers(0, n, size=(m, 2))
    w = rng.integers(1, 6, size=m).astype(np.float32) if ties else rng.uniform(0, 1, size=m).astype(np.float32)
    return _csr_from_edges(n, e, w)


def _kept_sorted(n, rp, col, w, thr):
    idx, d = [], []
    for v in range(n):
        for t in range(int(rp[v]), int(rp[v + 1])):
            u = int(col[t])
            if u < v and w[t] <= thr:
                idx.append(v * (v - 1) // 2 + u)
                d.append(w[t])
    idx = np.array(idx, np.uint64)
    d = np.array(d, np.float32)
    order = np.lexsort((-idx.astype(np.int64), d))     # d ascending, then idx descending
    return idx[order], d[order]


def _decode(idx):
    v = int((1 + math.isqrt(1 + 8 * int(idx))) // 2)
    while v * (v - 1) // 2 > idx:
        v -= 1
    while (v + 1) * v // 2 <= idx:
        v += 1
    return v, int(idx) - v * (v - 1) // 2


def _brute_apparent(n, rp, col, w, idx, d):
    nb = []
    for v in range(n):
        s = slice(int(rp[v]), int(rp[v + 1]))
        nb.append(dict(zip(col[s].tolist(), w[s].tolist())))
    out = np.zeros(idx.size, np.uint8)
    for e, (ix, dd) in enumerate(zip(idx.tolist(), d.tolist())):
        a, b = _decode(ix)
        common = sorted(set(nb[a]) & set(nb[b]), reverse=True)
        x = next((x for x in common if max(nb[a][x], nb[b][x]) <= dd), None)
        if x is None:
            continue
        v0, v1, v2 = sorted((a, b, x), reverse=True)
        facets = [(v1, v2), (v0, v2), (v0, v1)]          # ascending index order
        first = next(f for f in facets if nb[f[0]][f[1]] == dd)
        out[e] = first == (a, b)
    return out


@pytest.mark.parametrize("seed,n,m,ties,thr", [(1, 30, 150, True, math.inf), (2, 40, 300, True, 3.0),
                                               (3, 50, 400, False, math.inf), (4, 25, 300, True, 2.0),
                                               (5, 60, 90, False, 0.5)])
def test_sparse_edges_brute_force(gpu, seed, n, m, ties, thr):
    from turbotops_amd import sparse_edges
    rp, col, w = _random_graph(seed, n, m, ties)
    print(f"case sparse brute seed={seed} n={n} m={m} ties={ties}", flush=True)
    idx, d, ap = sparse_edges((n, rp, col, w), threshold=thr)
    want_idx, want_d = _kept_sorted(n, rp, col, w, thr)
    assert np.array_equal(idx, want_idx)
    assert np.array_equal(d.view(np.uint32), want_d.view(np.uint32))
    assert np.array_equal(ap, _brute_apparent(n, rp, col, w, idx, d))
    if ties:
        assert ap.any() and not ap.all()


@pytest.mark.parametrize("kind", ["knn", "radius"])
def test_sparse_edges_vs_oracle(gpu, restatement, kind):
    from oracle import Complex
    from turbotops_amd import knn_graph_exact, radius_graph, sparse_edges
    rng = np.random.default_rng(31 if kind == "knn" else 32)
    x = rng.uniform(size=(3000, 4)).astype(np.float32)
    g = knn_graph_exact(x, "l2", 8) if kind == "knn" else radius_graph(x, "l2", 0.22)
    print(f"case sparse vs oracle {kind} n=3000 nnz={g.row_ptr[-1]}", flush=True)
    idx, d, ap = sparse_edges(g)
    want_idx, want_d = _kept_sorted(g.n, g.row_ptr, g.col_idx, g.weights, math.inf)
    assert np.array_equal(idx, want_idx) and np.array_equal(d, want_d)
    cx = Complex(n=g.n, row_ptr=g.row_ptr, col=g.col_idx, w=g.weights)
    flags = restatement.apparent_cx(cx, 2, [int(v) for v in idx], d) & 1
    assert np.array_equal(ap, flags)
    assert 0 < ap.sum() < ap.size


def test_sparse_edges_degenerate_and_errors(gpu):
    from turbotops_amd import sparse_edges
    empty = (0, np.zeros(1, np.uint64), np.zeros(0, np.uint32), np.zeros(0, np.float32))
    idx, d, ap = sparse_edges(empty)
    assert idx.size == 0 and ap.size == 0
    lone = (3, np.zeros(4, np.uint64), np.zeros(0, np.uint32), np.zeros(0, np.float32))
    assert sparse_edges(lone)[0].size == 0
    rp, col, w = _csr_from_edges(3, [(0, 1), (1, 2)], [1.0, 2.0])
    assert sparse_edges((3, rp, col, w), apparent=False)[2] is None
    bad = col.copy()
    bad[1], bad[2] = bad[2], bad[1]                       # row 1 becomes [2, 0]
    with pytest.raises(ValueError, match="sorted"):
        sparse_edges((3, rp, bad, w))
    bad = col.copy()
    bad[0] = 7
    with pytest.raises(ValueError, match="out of range"):
        sparse_edges((3, rp, bad, w))
    with pytest.raises(ValueError, match="non-negative"):
        sparse_edges((3, rp, col, -w))
    bad = col.copy()
    bad[0] = 0                                            # row 0 = [0]: a self loop
    with pytest.raises(ValueError, match="self loop"):
        sparse_edges((3, rp, bad, w))
    with pytest.raises(ValueError, match="threshold"):
        sparse_edges((3, rp, col, w), threshold=float("nan"))
