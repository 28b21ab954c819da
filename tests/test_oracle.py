"""Pins the oracle (oracle/ttoracle.c) against everything except itself.

* The reference's own goldens (tests/test_cli.sh:56-94) and self-test
  invariants (src/turbotops.cpp:387-441).
* Golden fixtures written by the compiled reference (tests/golden/).
* The compiled reference itself (oracle/_ref), when present.
* Brute force and textbook algorithms for what a total order or a closed
  form fixes (kNN order, radius threshold, Kruskal's MST multiset, the
  apparent-pair definition).
* Planted mistakes: a perturbed oracle output must fail the same checks.
"""
import itertools
import math

import numpy as np
import pytest

from conftest import bits

METRICS = ["l1", "l2", "linf", "lp:1.5", "lp:3", "lp:4", "lp:7.5", "lp:50"]
SQUARE = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32)


# ---- closed forms from the reference's CLI goldens ---------------------------

@pytest.mark.parametrize("metric,diag", [("l2", "1.41421354"), ("l1", "2"), ("lp:3", "1.25992107"),
                                         ("linf", "1")])
def test_unit_square_goldens(restatement, metric, diag):
    # tests/test_cli.sh:56-94: sides 1, diagonal printed with %.9g
    mat = restatement.make_dense(SQUARE, metric)
    assert "%.9g" % mat[0, 2] == diag
    assert "%.9g" % mat[1, 3] == diag
    for i in range(4):
        assert mat[i, (i + 1) % 4] == 1.0
        assert mat[i, i] == 0.0


def test_l2_is_correctly_rounded_sqrt(restatement):
    # sqrt(2) rounded to float32 is 0x3fb504f3
    assert bits(restatement.dist("l2", [0, 0], [1, 1])) == 0x3FB504F3


# ---- the reference's self-test invariants (src/turbotops.cpp:387-416) -------

@pytest.mark.parametrize("dim", [3, 17])
def test_metric_axioms(restatement, dim):
    rng = np.random.default_rng(7)
    for p in [1.0, 1.5, 2.0, 3.0, 7.5, 50.0, math.inf]:
        m = "lp:inf" if math.isinf(p) else f"lp:{p}"
        for _ in range(60):
            a, b, c = rng.normal(scale=3.0, size=(3, dim)).astype(np.float32)
            assert restatement.dist(m, a, b) == restatement.dist(m, b, a)
            assert restatement.dist(m, a, a) == 0.0
            assert restatement.dist(m, a, c) <= (restatement.dist(m, a, b) + restatement.dist(m, b, c)) * (1 + 1e-5)


def test_large_p_and_one_dim(restatement):
    a = np.zeros(9, np.float32)
    b = (1000.0 * np.arange(1, 10)).astype(np.float32)
    d50, dinf = restatement.dist("lp:50", a, b), restatement.dist("linf", a, b)
    assert math.isfinite(d50) and dinf <= d50 <= dinf * 1.05
    assert abs(restatement.dist("lp:3", a[:1], b[:1]) - 1000.0) < 1e-3


def test_metric_parse(restatement):
    assert restatement.parse("l2") == (1, 2.0)
    assert restatement.parse("euclidean") == (1, 2.0)
    assert restatement.parse("manhattan") == (0, 1.0)
    assert restatement.parse("chebyshev")[0] == 2
    assert restatement.parse("lp:inf")[0] == 2
    assert restatement.parse("lp:2") == (1, 2.0)
    assert restatement.parse("lp:1") == (0, 1.0)
    assert restatement.parse("lp:1.5") == (3, 1.5)
    for bad in ["lp:0.5", "cosine", "lp:", "lp:nan", "lp:2x"]:
        with pytest.raises(ValueError):
            restatement.parse(bad)


# ---- golden fixtures written by the compiled reference -----------------------

@pytest.mark.parametrize("metric", METRICS)
def test_metric_pairs_golden(restatement, golden, metric):
    g = golden("metric_pairs.npz")
    a, b, dims = g["a"], g["b"], g["dims"]
    got = np.array([restatement.dist(metric, a[t, :dims[t]], b[t, :dims[t]]) for t in range(len(dims))],
                   np.float32)
    assert np.array_equal(bits(got), bits(g["d_" + metric]))


def _cloud_names(g):
    return sorted(k[2:] for k in g.files if k.startswith("x_"))


@pytest.mark.parametrize("metric", METRICS)
def test_dense_golden(restatement, golden, metric):
    g = golden("dense_small.npz")
    for name in _cloud_names(g):
        x = g["x_" + name]
        n = x.shape[0]
        full = restatement.make_dense(x, metric, True)
        assert np.array_equal(bits(full), bits(g[f"full_{name}_{metric}"])), name
        low = restatement.make_dense(x, metric, False)
        assert np.array_equal(bits(low), bits(g[f"lower_{name}_{metric}"])), name
        r = restatement.enclosing_radius(full, n, True)
        assert bits(r) == bits(g[f"radius_{name}_{metric}"])
        assert bits(restatement.enclosing_radius(low, n, False)) == bits(r)


@pytest.mark.parametrize("metric", METRICS)
def test_graphs_golden(restatement, golden, metric):
    g = golden("graphs_small.npz")
    d = golden("dense_small.npz")
    for name in _cloud_names(g):
        x = g["x_" + name]
        tau = float(np.float32(0.6) * d[f"radius_{name}_{metric}"])
        rp, col, w = restatement.radius_graph(x, metric, tau)
        assert np.array_equal(rp, g[f"rg_rp_{name}_{metric}"])
        assert np.array_equal(col, g[f"rg_col_{name}_{metric}"])
        assert np.array_equal(bits(w), bits(g[f"rg_w_{name}_{metric}"]))
        ids = restatement.knn(x, metric, 7)
        assert np.array_equal(ids, g[f"knn_{name}_{metric}"])
        rp, col, w = restatement.knn_graph(x, metric, ids)
        assert np.array_equal(rp, g[f"kg_rp_{name}_{metric}"])
        assert np.array_equal(col, g[f"kg_col_{name}_{metric}"])
        assert np.array_equal(bits(w), bits(g[f"kg_w_{name}_{metric}"]))
        deaths = restatement.mst_deaths(x, metric)
        ref_deaths = g[f"mst_{name}_{metric}"]
        assert np.array_equal(np.sort(ref_deaths[np.isfinite(ref_deaths)]), deaths[deaths > 0])
        mat = restatement.make_dense(x, metric, True)
        thr = restatement.enclosing_radius(mat, x.shape[0], True)
        assert bits(thr) == bits(g[f"edges_thr_{name}_{metric}"])
        idx, dd = restatement.dense_edges(mat, thr)
        assert np.array_equal(idx, g[f"edges_idx_{name}_{metric}"])
        assert np.array_equal(bits(dd), bits(g[f"edges_d_{name}_{metric}"]))
        fl = restatement.apparent_edges(mat, idx, dd)
        assert np.array_equal(fl, g[f"edges_apparent_{name}_{metric}"])


# ---- live comparison with the compiled reference -----------------------------

@pytest.mark.parametrize("metric", METRICS)
@pytest.mark.parametrize("shape", [(45, 3), (33, 16), (21, 1), (40, 9)])
def test_restatement_vs_reference(restatement, reference, metric, shape):
    rng = np.random.default_rng(hash((metric, shape)) % 2**32)
    x = rng.normal(size=shape).astype(np.float32)
    n = shape[0]
    full = restatement.make_dense(x, metric, True)
    ok, rfull = reference.make_dense(x, metric)
    assert ok and np.array_equal(bits(full), bits(rfull))
    assert bits(restatement.enclosing_radius(full, n)) == bits(reference.enclosing_radius(rfull, n))
    idx, d, ap, thr = reference.dense_edges(x, metric)
    i2, d2 = restatement.dense_edges(full, thr)
    assert np.array_equal(idx, i2) and np.array_equal(bits(d), bits(d2))
    assert np.array_equal(ap, restatement.apparent_edges(full, i2, d2))
    for a, b in zip(restatement.radius_graph(x, metric, 1.1), reference.radius_graph(x, metric, 1.1)):
        assert np.array_equal(a, b)
    assert np.array_equal(restatement.knn(x, metric, 5), reference.knn(x, metric, 5))


# ---- brute force / textbook pins ------------------------------------------------

def test_knn_matches_lexicographic_sort(restatement):
    rng = np.random.default_rng(3)
    x = rng.integers(0, 3, size=(40, 2)).astype(np.float32)   # many ties
    mat = restatement.make_dense(x, "l1")
    ids = restatement.knn(x, "l1", 6)
    for i in range(40):
        order = sorted((mat[i, j], j) for j in range(40) if j != i)
        assert [j for _, j in order[:6]] == list(ids[i])


def test_radius_graph_is_threshold(restatement):
    rng = np.random.default_rng(4)
    x = rng.normal(size=(50, 4)).astype(np.float32)
    mat = restatement.make_dense(x, "l2")
    rp, col, w = restatement.radius_graph(x, "l2", 1.3)
    for i in range(50):
        js = [j for j in range(50) if j != i and mat[i, j] <= np.float32(1.3)]
        assert list(col[rp[i]:rp[i + 1]]) == js
        assert np.array_equal(bits(w[rp[i]:rp[i + 1]]), bits(mat[i, js]))


def _kruskal(mat):
    n = mat.shape[0]
    edges = sorted((mat[i, j], i, j) for i in range(n) for j in range(i))
    parent = list(range(n))

    def find(a):
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a
    out = []
    for w, i, j in edges:
        a, b = find(i), find(j)
        if a != b:
            parent[a] = b
            out.append(w)
    return np.array(out, np.float32)


@pytest.mark.parametrize("metric", ["l2", "linf", "lp:1.5"])
def test_mst_is_kruskal_multiset(restatement, metric):
    rng = np.random.default_rng(5)
    x = rng.normal(size=(60, 3)).astype(np.float32)
    deaths = restatement.mst_deaths(x, metric)
    assert np.array_equal(bits(deaths), bits(_kruskal(restatement.make_dense(x, metric))))


def _apparent_brute(mat, i, j, d):
    """Definition (rips.hpp:585-649) restated from scratch on one edge."""
    n = mat.shape[0]
    ws = [w for w in range(n) if w not in (i, j) and max(mat[i, w], mat[j, w]) <= d]
    if not ws:
        return False
    w = max(ws)
    tri = sorted((i, j, w), reverse=True)
    # facets in ascending combinatorial index: drop tri[0], then tri[1], then tri[2]
    for drop in range(3):
        rest = [v for t, v in enumerate(tri) if t != drop]
        if mat[rest[0], rest[1]] == d:
            return sorted(rest, reverse=True) == sorted((i, j), reverse=True)
    return False


@pytest.mark.parametrize("kind", ["gauss", "grid"])
def test_apparent_edges_brute_force(restatement, kind):
    rng = np.random.default_rng(6)
    x = (rng.normal(size=(35, 3)) if kind == "gauss" else rng.integers(0, 3, size=(35, 3))).astype(np.float32)
    mat = restatement.make_dense(x, "l2")
    thr = restatement.enclosing_radius(mat, 35)
    idx, d = restatement.dense_edges(mat, thr)
    flags = restatement.apparent_edges(mat, idx, d)
    for t in range(len(idx)):
        i, j = restatement.decode(35, 2, int(idx[t]))
        assert bool(flags[t]) == _apparent_brute(mat, int(i), int(j), d[t])


def test_edge_order_is_total(restatement):
    rng = np.random.default_rng(8)
    x = rng.integers(0, 3, size=(30, 2)).astype(np.float32)
    mat = restatement.make_dense(x, "l1")
    idx, d = restatement.dense_edges(mat, np.inf)
    keys = list(zip(d.tolist(), (-idx.astype(np.int64)).tolist()))
    assert keys == sorted(keys) and len(set(keys)) == len(keys)
    assert len(idx) == 30 * 29 // 2


def test_decode_roundtrip(restatement):
    from math import comb
    for vs in itertools.combinations(range(12), 3):
        v = sorted(vs, reverse=True)
        idx = sum(comb(v[t], 3 - t) for t in range(3))
        assert list(restatement.decode(12, 3, idx)) == v
    big = [199999, 123456, 77, 3]
    idx = sum(comb(big[t], 4 - t) for t in range(4))
    assert idx > 2**64   # u128 regime of C(2e5, 4)
    assert list(restatement.decode(200000, 4, idx)) == big


# ---- planted mistakes: the checks above must catch them ---------------------

def test_planted_mistakes_are_caught(restatement, golden):
    g = golden("dense_small.npz")
    x = g["x_gauss37x3"]
    want = g["full_gauss37x3_l2"]
    got = restatement.make_dense(x, "l2")
    # a dropped term, a transposed operand, a wrong rounding: each must differ
    dropped = restatement.make_dense(x[:, :2], "l2")
    assert not np.array_equal(bits(dropped), bits(want))
    assert not np.array_equal(bits(restatement.make_dense(x, "l1")), bits(want))
    nofma = np.sqrt(((x[:, None, :] - x[None, :, :]) ** 2).sum(-1, dtype=np.float32)).astype(np.float32)
    assert np.array_equal(bits(got), bits(want))
    assert (bits(nofma) != bits(want)).sum() > 0   # a non-fma sum is caught bitwise


# ---- the CPU reference gradient (oracle/grad.py) -----------------------------

GRAD_METRICS = ["l2", "l1", "linf", "lp:1.5", "lp:3", "lp:7.5"]


def _grad_case(seed, n=40, m=60, dim=4):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(n, dim)).astype(np.float32)
    e = rng.integers(0, n, size=(m, 2))
    e = e[e[:, 0] != e[:, 1]]
    e = np.concatenate([e, e[:3], e[:2, ::-1]])      # repeated and reversed edges
    g = rng.normal(size=e.shape[0])
    return x, e, g


def _fd64(x, e, g, metric, h=1e-6):
    from oracle.grad import dist64
    x64 = x.astype(np.float64)
    out = np.zeros_like(x64)
    loss = lambda z: float((g * dist64(z, e[:, 0], e[:, 1], metric)).sum())
    for i in range(x.shape[0]):
        for k in range(x.shape[1]):
            zp, zm = x64.copy(), x64.copy()
            zp[i, k] += h
            zm[i, k] -= h
            out[i, k] = (loss(zp) - loss(zm)) / (2 * h)
    return out


@pytest.mark.parametrize("metric", GRAD_METRICS)
def test_grad_oracle_central_differences(metric):
    """Central differences in float64 of L = sum g_e d(x_a, x_b) pin every
    metric's derivative (generic points: no ties, no zero distances)."""
    from oracle.grad import edge_grad64
    x, e, g = _grad_case(21)
    want = _fd64(x, e, g, metric)
    got = edge_grad64(x, e, g, metric)
    np.testing.assert_allclose(got, want, rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("metric,order", [("l2", 2), ("l1", 1), ("linf", math.inf)])
def test_grad_oracle_matches_torch_autograd(metric, order):
    import torch
    from oracle.grad import edge_grad64
    x, e, g = _grad_case(22)
    xt = torch.tensor(x, dtype=torch.float64, requires_grad=True)
    d = torch.linalg.vector_norm(xt[e[:, 0]] - xt[e[:, 1]], ord=order, dim=1)
    (torch.tensor(g) * d).sum().backward()
    np.testing.assert_allclose(edge_grad64(x, e, g, metric), xt.grad.numpy(), rtol=1e-12, atol=1e-12)


def test_grad_oracle_degenerate_and_planted():
    """A zero-length edge contributes nothing; a planted sign flip, a dropped
    endpoint or a wrong Lp exponent each fail the difference check."""
    from oracle.grad import edge_grad64
    x, e, g = _grad_case(23)
    x[1] = x[0]
    z = edge_grad64(x, np.array([[0, 1], [1, 1]]), np.array([2.0, 3.0]), "l2")
    assert not z.any()
    want = _fd64(x, e[e[:, 0] + e[:, 1] > 1], g[e[:, 0] + e[:, 1] > 1], "lp:3")
    ok = edge_grad64(x, e[e[:, 0] + e[:, 1] > 1], g[e[:, 0] + e[:, 1] > 1], "lp:3")
    np.testing.assert_allclose(ok, want, rtol=1e-6, atol=1e-6)
    keep = e[:, 0] + e[:, 1] > 1
    ek, gk = e[keep], g[keep]
    flipped = edge_grad64(x, ek[:, ::-1], -gk, "lp:3")          # sign error
    assert not np.allclose(flipped, want, rtol=1e-6, atol=1e-6)
    dropped = np.zeros_like(ok)                                  # endpoint b forgotten
    for (a, b), gv in zip(ek, gk):
        one = edge_grad64(x, np.array([[a, b]]), np.array([gv]), "lp:3")
        dropped[a] += one[a] if a != b else 0
    assert not np.allclose(dropped, want, rtol=1e-6, atol=1e-6)
    wrong_p = edge_grad64(x, ek, gk, "lp:2.9")
    assert not np.allclose(wrong_p, want, rtol=1e-6, atol=1e-6)
