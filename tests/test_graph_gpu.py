"""GPU parity of the neighbourhood-graph and H0 kernels against the oracle:
radius_graph (graph.hpp:170-201), knn_lists_exact (:211-225), knn_graph
(:286-304), mst::h0 (mst.hpp:25-99). Index work bit-exact; weights bit-exact
for L1, L2 and L-inf, within 2 ulp for Lp."""
import numpy as np
import pytest

from conftest import bits, ulp_diff

pytestmark = pytest.mark.gpu

METRICS = ["l1", "l2", "linf", "lp:1.5", "lp:4"]


def _cloud(n, d, seed, kind="gauss"):
    rng = np.random.default_rng(seed)
    if kind == "grid":
        return rng.integers(0, 4, size=(n, d)).astype(np.float32)
    if kind == "unif":
        return rng.uniform(size=(n, d)).astype(np.float32)
    return rng.normal(size=(n, d)).astype(np.float32)


def _same_w(a, b, metric):
    if metric in ("l1", "l2", "linf"):
        return np.array_equal(bits(a), bits(b))
    return a.size == b.size and (a.size == 0 or ulp_diff(a, b).max() <= 2)


@pytest.mark.parametrize("metric", METRICS)
def test_reference_selftest_case(gpu, restatement, metric):
    """The reference's own check (src/turbotops.cpp:418-441): n=300, d=4
    gaussian, exact kNN k=7 and radius graph tau=1.2 against brute force."""
    from turbotops_amd import knn_lists_exact, radius_graph
    x = _cloud(300, 4, 11)
    print(f"case selftest metric={metric}", flush=True)
    ids = knn_lists_exact(x, metric, 7)
    if metric in ("l1", "l2", "linf"):
        assert np.array_equal(ids, restatement.knn(x, metric, 7))
    else:
        _lp_lists_agree(restatement, x, metric, ids)
    g = radius_graph(x, metric, 1.2)
    rp, col, w = restatement.radius_graph(x, metric, 1.2)
    if metric in ("l1", "l2", "linf"):
        assert g.valid() and np.array_equal(g.row_ptr, rp) and np.array_equal(g.col_idx, col)
        assert _same_w(g.weights, w, metric)
    else:
        assert g.valid()
        _lp_graphs_agree(g, rp, col, w, 1.2)
    mat = restatement.make_dense(x, metric)
    assert g.m == int((mat[np.triu_indices(300, 1)] <= np.float32(1.2)).sum())


def _lp_lists_agree(restatement, x, metric, ids):
    """Lp distances may differ from glibc powf by an ulp (2 at most): each
    list must hold the oracle's neighbours wherever its k-th and (k+1)-th
    distances are more than 4 ulp apart, and then in the same order wherever
    consecutive distances are."""
    n, k = ids.shape
    mat = restatement.make_dense(x, metric)
    want = restatement.knn(x, metric, k)
    checked = 0
    for i in range(n):
        row = np.delete(mat[i], i)
        srt = np.sort(row)
        if k < n - 1 and ulp_diff(srt[k - 1], srt[k]) <= 4:
            continue
        assert set(ids[i].tolist()) == set(want[i].tolist()), i
        dw = mat[i, want[i]]
        sep = np.r_[True, ulp_diff(dw[1:], dw[:-1]) > 4] & np.r_[ulp_diff(dw[1:], dw[:-1]) > 4, True]
        assert np.array_equal(ids[i][sep], want[i][sep]), i
        checked += 1
    assert checked > n // 2


def _lp_graphs_agree(g, rp, col, w, tau):
    """Every edge whose oracle weight is more than 2 ulp from tau is in both
    graphs with weights within 2 ulp; edges within 2 ulp of tau may differ."""
    def edges(rp_, col_, w_):
        out = {}
        for v in range(len(rp_) - 1):
            for t in range(int(rp_[v]), int(rp_[v + 1])):
                out[(v, int(col_[t]))] = np.float32(w_[t])
        return out
    got = edges(g.row_ptr, g.col_idx, g.weights)
    want = edges(rp, col, w)
    t32 = np.float32(tau)
    for e, wv in want.items():
        if ulp_diff(wv, t32) > 2:
            assert e in got and ulp_diff(got[e], wv) <= 2, e
    for e, wv in got.items():
        if ulp_diff(wv, t32) > 2:
            assert e in want, e
    assert len(want) > 0


@pytest.mark.parametrize("metric", ["l1", "l2", "linf"])
@pytest.mark.parametrize("n,d,kind,k", [(2, 3, "gauss", 5), (50, 3, "grid", 7), (257, 2, "grid", 8),
                                        (300, 8, "unif", 15), (130, 16, "gauss", 10), (90, 5, "gauss", 33),
                                        (200, 3, "gauss", 64), (70, 17, "gauss", 4), (1, 3, "gauss", 3)])
def test_knn_and_knn_graph(gpu, restatement, metric, n, d, kind, k):
    from turbotops_amd import knn_graph, knn_graph_exact, knn_lists_exact
    x = _cloud(n, d, n + d, kind)
    print(f"case knn metric={metric} n={n} d={d} k={k}", flush=True)
    ids = knn_lists_exact(x, metric, k)
    want = restatement.knn(x, metric, k)
    assert ids.shape == want.shape and np.array_equal(ids, want)
    g = knn_graph(x, metric, ids)
    rp, col, w = restatement.knn_graph(x, metric, want)
    assert np.array_equal(g.row_ptr, rp) and np.array_equal(g.col_idx, col)
    assert np.array_equal(bits(g.weights), bits(w))
    g2 = knn_graph_exact(x, metric, k)
    assert np.array_equal(g2.row_ptr, rp) and np.array_equal(g2.col_idx, col)
    assert g.valid()


@pytest.mark.parametrize("metric,n,d,kind,k", [("l2", 8192, 3, "grid", 15), ("l1", 8300, 8, "unif", 64),
                                                ("linf", 9000, 2, "grid", 9), ("l2", 12300, 4, "gauss", 1)])
def test_knn_candidate_chunks(gpu, restatement, metric, n, d, kind, k):
    """Clouds large enough that the candidates are split into chunks whose
    partial lists are merged; grid clouds put equal distances in different
    chunks, where the smaller index must win."""
    from turbotops_amd import knn_lists_exact
    x = _cloud(n, d, n + 3 * d, kind)
    print(f"case knn chunks metric={metric} n={n} d={d} k={k}", flush=True)
    ids = knn_lists_exact(x, metric, k)
    want = restatement.knn(x, metric, k)
    assert ids.shape == want.shape and np.array_equal(ids, want)


@pytest.mark.parametrize("metric", METRICS)
@pytest.mark.parametrize("shape,k", [("clusters", 15), ("duplicates", 15), ("duplicates", 64), ("line", 30),
                                     ("aniso", 15)])
def test_knn_seeded_scan_cases(gpu, restatement, metric, shape, k):
    """The scan walks a spatial order and admits only candidates within each
    query's seed bound (its k-th distance among its 4096 neighbours in that
    order). Clusters far apart put several clusters in one seed range;
    duplicates make the bound 0 with the k best decided by id among equal
    distances; a segment has boxes of zero width; the 5-D cloud runs the
    runtime-dimension scan. n = 9000 spans three seed ranges and chunks."""
    from turbotops_amd import knn_lists_exact
    x = _stress_cloud(shape, 9000, 43)
    print(f"case knn-seeded metric={metric} shape={shape} k={k}", flush=True)
    ids = knn_lists_exact(x, metric, k)
    want = restatement.knn(x, metric, k)
    if metric.startswith("lp"):
        # Lp distances may sit within 2 ulp of the oracle's, so lists may swap
        # neighbours at near-equal distances: every list must reach the same
        # distances as the oracle's, rank by rank, recomputed in float64
        p = float(metric.split(":")[1])
        assert ids.shape == want.shape
        xd = x.astype(np.float64)

        def d64(lists):
            return np.sort((np.abs(xd[:, None, :] - xd[lists]) ** p).sum(-1) ** (1 / p), axis=1)
        assert np.allclose(d64(ids), d64(want), rtol=1e-6, atol=1e-30)
    else:
        assert ids.shape == want.shape and np.array_equal(ids, want)


@pytest.mark.parametrize("metric", ["l1", "l2", "linf"])
@pytest.mark.parametrize("n,d,kind,k", [(300, 3, "grid", 65), (700, 4, "gauss", 100), (1500, 8, "unif", 500),
                                        (120, 2, "grid", 119), (2000, 3, "grid", 300)])
def test_knn_large_k(gpu, restatement, metric, n, d, kind, k):
    """k > 64 (knn_lists_exact clamps only to n - 1, graph.hpp:214): the
    selection path, ties at the k-th distance included (integer grids), the
    whole row at k = n - 1; ids bit-exact, and the kNN graph."""
    from turbotops_amd import knn_graph, knn_graph_exact, knn_lists_exact
    x = _cloud(n, d, n + 5 * d, kind)
    print(f"case knn large k metric={metric} n={n} d={d} k={k}", flush=True)
    ids = knn_lists_exact(x, metric, k)
    want = restatement.knn(x, metric, k)
    assert ids.shape == want.shape == (n, min(k, n - 1)) and np.array_equal(ids, want)
    g = knn_graph(x, metric, ids)
    rp, col, w = restatement.knn_graph(x, metric, want)
    assert np.array_equal(g.row_ptr, rp) and np.array_equal(g.col_idx, col)
    assert np.array_equal(bits(g.weights), bits(w))
    g2 = knn_graph_exact(x, metric, k)
    assert np.array_equal(g2.row_ptr, rp) and np.array_equal(g2.col_idx, col)


def test_knn_large_k_lp(gpu, restatement):
    from turbotops_amd import knn_lists_exact
    x = _cloud(400, 5, 77)
    _lp_lists_agree(restatement, x, "lp:1.5", knn_lists_exact(x, "lp:1.5", 90))


@pytest.mark.parametrize("metric", ["l2", "l1"])
def test_persistence_knn_k100_vs_reference(gpu, reference, metric):
    """persistence(x, graph="knn", k=100), which the reference accepts for any
    k (graph.hpp:211-225), bar for bar against the compiled reference."""
    from turbotops_amd import persistence
    x = _cloud(300, 3, 5)
    dg, _ = persistence(x, maxdim=1, metric=metric, graph="knn", k=100)
    (dims, b, e), _ = reference.persistence(x, maxdim=1, metric=metric, graph="knn", k=100)
    got = sorted((kk, float(bb), float(ee)) for kk, a in enumerate(dg) for bb, ee in np.asarray(a).tolist())
    assert got == sorted(zip(dims.tolist(), b.tolist(), e.tolist()))


@pytest.mark.parametrize("metric", ["lp:1.5", "lp:3"])
def test_knn_lp(gpu, restatement, metric):
    """Lp distances may differ from glibc powf by an ulp; lists must agree
    wherever the oracle's k-th and (k+1)-th distances are separated by more."""
    from turbotops_amd import knn_lists_exact
    x = _cloud(150, 5, 3)
    ids = knn_lists_exact(x, metric, 9)
    mat = restatement.make_dense(x, metric)
    want = restatement.knn(x, metric, 9)
    for i in range(150):
        row = np.sort(np.delete(mat[i], i))
        if ulp_diff(row[8], row[9]) > 4:
            assert set(ids[i]) == set(want[i])


@pytest.mark.parametrize("metric", METRICS)
@pytest.mark.parametrize("n,d,kind,tau", [(2, 3, "gauss", 10.0), (64, 3, "grid", 1.0), (65, 3, "grid", 2.0),
                                          (300, 8, "unif", 0.6), (500, 2, "gauss", 0.2), (120, 17, "gauss", 4.0),
                                          (150, 3, "gauss", 1e30), (40, 3, "gauss", 0.0)])
def test_radius_graph(gpu, restatement, metric, n, d, kind, tau):
    from turbotops_amd import radius_graph
    x = _cloud(n, d, 7 * n + d, kind)
    print(f"case radius metric={metric} n={n} d={d} tau={tau}", flush=True)
    g = radius_graph(x, metric, tau)
    rp, col, w = restatement.radius_graph(x, metric, tau)
    if metric in ("l1", "l2", "linf"):
        assert np.array_equal(g.row_ptr, rp) and np.array_equal(g.col_idx, col)
        assert np.array_equal(bits(g.weights), bits(w))
    else:   # an ulp at the threshold may move an edge; away from it the graphs agree
        assert abs(int(g.row_ptr[-1]) - int(rp[-1])) <= 2
        if rp[-1]:
            _lp_graphs_agree(g, rp, col, w, tau)
    assert g.valid()


@pytest.mark.parametrize("metric", METRICS)
@pytest.mark.parametrize("shape,tau", [("clusters", 0.03), ("clusters", 40.0), ("duplicates", 0.0),
                                       ("duplicates", 0.5), ("line", 1e-3), ("aniso", 1.5)])
def test_radius_graph_pruned_tiles(gpu, restatement, shape, tau, metric):
    """Tiles whose blocks' boxes lie farther apart than tau are skipped
    whole; these inputs put tau at and around the box gaps: tight clusters
    (tau inside a cluster, and tau spanning some clusters), duplicates (tau =
    0 keeps exactly the coincident pairs), a segment (boxes of zero width)
    and an anisotropic 5-D cloud (runtime dimension), n = 3000."""
    from turbotops_amd import radius_graph
    x = _stress_cloud(shape, 3000, 47)
    print(f"case radius-pruned metric={metric} shape={shape} tau={tau}", flush=True)
    g = radius_graph(x, metric, tau)
    rp, col, w = restatement.radius_graph(x, metric, tau)
    if metric in ("l1", "l2", "linf"):
        assert np.array_equal(g.row_ptr, rp) and np.array_equal(g.col_idx, col)
        assert np.array_equal(bits(g.weights), bits(w))
    else:
        assert abs(int(g.row_ptr[-1]) - int(rp[-1])) <= 2
        if rp[-1]:
            _lp_graphs_agree(g, rp, col, w, tau)
    assert g.valid()


def test_radius_graph_large(gpu, restatement):
    """Edge count beyond the first buffer guess (complete graph, n = 2000):
    the retry path, then CSR invariants and sampled rows."""
    from turbotops_amd import radius_graph
    x = _cloud(2000, 3, 5)
    g = radius_graph(x, "l2", 1e30)
    assert g.m == 2000 * 1999 // 2
    rp = g.row_ptr
    assert np.all(np.diff(rp.astype(np.int64)) == 1999)
    for v in (0, 1, 777, 1999):
        cols = g.col_idx[rp[v]:rp[v + 1]]
        assert np.array_equal(cols, np.delete(np.arange(2000), v))
        want = np.array([restatement.dist("l2", x[v], x[u]) for u in cols], np.float32)
        assert np.array_equal(bits(g.weights[rp[v]:rp[v + 1]]), bits(want))


@pytest.mark.parametrize("metric", METRICS)
@pytest.mark.parametrize("n,d,kind", [(2, 3, "gauss"), (3, 2, "grid"), (100, 3, "gauss"), (257, 2, "grid"),
                                      (600, 8, "unif"), (333, 17, "gauss"), (1000, 3, "grid")])
def test_mst_deaths(gpu, restatement, metric, n, d, kind):
    from turbotops_amd import mst_deaths
    x = _cloud(n, d, 3 * n + d, kind)
    print(f"case mst metric={metric} n={n} d={d} {kind}", flush=True)
    got = mst_deaths(x, metric)
    want = restatement.mst_deaths(x, metric)
    if metric in ("l1", "l2", "linf"):
        assert np.array_equal(bits(got), bits(want))
    else:
        assert ulp_diff(got, want).max() <= 2


def _stress_cloud(shape, n, seed):
    rng = np.random.default_rng(seed)
    if shape == "clusters":    # tight clusters far apart, stored interleaved
        centres = rng.uniform(-50, 50, size=(7, 3))
        return (centres[rng.integers(0, 7, size=n)] + rng.normal(scale=0.01, size=(n, 3))).astype(np.float32)
    if shape == "duplicates":  # 40 distinct points, each repeated
        pts = rng.normal(size=(40, 2))
        return pts[rng.integers(0, 40, size=n)].astype(np.float32)
    if shape == "line":        # a segment in R^3: boxes of zero width in two coordinates
        t = rng.uniform(size=(n, 1))
        return np.hstack([t, np.full((n, 1), 0.25), np.full((n, 1), -1.0)]).astype(np.float32)
    return (rng.uniform(size=(n, 5)) * np.array([1, 10, 100, 0.1, 1])).astype(np.float32)   # "aniso", runtime dim


@pytest.mark.parametrize("metric", METRICS)
@pytest.mark.parametrize("shape", ["clusters", "duplicates", "line", "aniso"])
def test_mst_pruned_search_cases(gpu, restatement, metric, shape):
    """The Boruvka search skips blocks and super-blocks by their boxes and
    component tags; these inputs stress the skips over several 4096-point
    super-blocks (n = 9000): clusters far apart (late rounds search across
    empty space, whole blocks in one component), many duplicates (keys of 0,
    ties), a segment (boxes of zero width), and an anisotropic 5-D cloud (the
    runtime-dimension kernel)."""
    from turbotops_amd import mst_deaths
    x = _stress_cloud(shape, 9000, 41)
    print(f"case mst-stress metric={metric} shape={shape}", flush=True)
    got = mst_deaths(x, metric)
    want = restatement.mst_deaths(x, metric)
    if metric in ("l1", "l2", "linf"):
        assert np.array_equal(bits(got), bits(want))
    else:
        assert ulp_diff(got, want).max() <= 2


def test_mst_h0_bars_and_threshold(gpu, reference):
    """Bars against the compiled reference's mst::h0, with and without a cut
    (src/turbotops.cpp:529-547)."""
    from turbotops_amd import mst_h0
    x = _cloud(400, 3, 13)
    for thr in [np.inf, 0.1, 0.05, 0.0]:
        b, d = mst_h0(x, "l2", thr)
        _, rb, rd = reference.mst_h0(x, "l2", thr)
        assert np.array_equal(bits(d), bits(rd)) and np.array_equal(b, rb)


def test_mst_rejects_non_finite(gpu):
    from turbotops_amd import mst_deaths
    x = _cloud(10, 3, 1)
    x[3, 1] = np.nan
    with pytest.raises(ValueError, match="mst: non-finite coordinate"):
        mst_deaths(x, "l2")


def test_mst_full_size_property(gpu, restatement):
    """n = 100000 uniform in [0,1]^3: n - 1 sorted non-negative deaths, and
    an exact oracle match on a 6000-point prefix of the same cloud."""
    from turbotops_amd import mst_deaths
    rng = np.random.default_rng(4)
    x = rng.uniform(size=(100000, 3)).astype(np.float32)
    d = mst_deaths(x, "l2")
    assert d.size == 99999 and np.all(np.diff(d) >= 0) and d[0] >= 0
    small = x[:6000]
    assert np.array_equal(bits(mst_deaths(small, "l2")), bits(restatement.mst_deaths(small, "l2")))
