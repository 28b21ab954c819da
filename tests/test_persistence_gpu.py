"""End-to-end diagrams: GPU front end + host reducer (turbotops_amd.persistence)
against the reference's goldens (tests/test_cli.sh:40-117), its Python parity
cases (tests/test_python.py:35-74), its self-test consistency checks
(src/turbotops.cpp:443-476, :529-619) and the compiled reference's own
tops::persistence (oracle/_ref), bar for bar."""
import math

import numpy as np
import pytest

from conftest import ulp_diff

pytestmark = pytest.mark.gpu

SQUARE = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32)
TRIANGLE = np.array([[0, 0], [1, 0], [0, 1]], np.float32)
SQ_L2 = ["0 0 1", "0 0 1", "0 0 1", "0 0 inf", "1 1 1.41421354"]
H0SQ = ["0 0 1", "0 0 1", "0 0 1", "0 0 inf"]
SQ_L1 = ["0 0 1", "0 0 1", "0 0 1", "0 0 inf", "1 1 2"]
SQ_LP3 = ["0 0 1", "0 0 1", "0 0 1", "0 0 inf", "1 1 1.25992107"]
SQ_THR = ["0 0 1", "0 0 1", "0 0 1", "0 0 inf", "1 1 inf"]
TRI = ["0 0 1", "0 0 1", "0 0 inf"]


def canon(dgms):
    """The CLI's canonical text (src/turbotops.cpp:208-230)."""
    bars = sorted((d, float(b), float(e)) for d, a in enumerate(dgms) for b, e in np.asarray(a))
    return [f"{d} {b:.9g} inf" if math.isinf(e) else f"{d} {b:.9g} {e:.9g}" for d, b, e in bars]


def ref_canon(res):
    dims, b, e = res
    return canon([np.stack([b[dims == d], e[dims == d]], 1) for d in range(int(dims.max()) + 1 if dims.size else 0)])


def bars_close(a, b, rtol):
    """same_bars (src/turbotops.cpp:288-326): zero-length bars dropped, sorted
    per dimension, each value within rtol."""
    A = sorted(canon_tuples(a))
    B = sorted(canon_tuples(b))
    if len(A) != len(B):
        return False
    for (d1, b1, e1), (d2, b2, e2) in zip(A, B):
        if d1 != d2 or math.isinf(e1) != math.isinf(e2):
            return False
        for u, v in ((b1, b2), (e1, e2)):
            if math.isfinite(u) and abs(u - v) > rtol * max(1.0, abs(v)):
                return False
    return True


def canon_tuples(dgms):
    return [(d, float(b), float(e)) for d, a in enumerate(dgms) for b, e in np.asarray(a) if e > b]


def ref_dgms(res, maxdim):
    dims, b, e = res
    return [np.stack([b[dims == d], e[dims == d]], 1) for d in range(maxdim + 1)]


# ---- the reference's CLI goldens ----------------------------------------------

@pytest.mark.parametrize("kw,gold", [
    (dict(maxdim=1), SQ_L2), (dict(maxdim=1, metric="l1"), SQ_L1), (dict(maxdim=1, metric="linf"), H0SQ),
    (dict(maxdim=1, metric="lp:3"), SQ_LP3), (dict(maxdim=1, threshold=1.0), SQ_THR),
    (dict(maxdim=1, graph="knn", k=2), SQ_THR), (dict(maxdim=2), SQ_L2)])
def test_cli_goldens_square(gpu, kw, gold):
    from turbotops_amd import persistence
    dgms, info = persistence(SQUARE, **kw)
    assert canon(dgms) == gold
    assert len(dgms) == kw.get("maxdim", 1) + 1


def test_cli_golden_triangle_and_edges(gpu):
    import scipy.sparse as sp
    from turbotops_amd import persistence
    dgms, _ = persistence(TRIANGLE)
    assert canon(dgms) == TRI
    # the square as a weighted edge list with a duplicate edge (smaller weight
    # kept, tests/test_cli.sh:49-50); CSR arrays built directly, as scipy's
    # COO conversion would sum the duplicate instead
    edges = [(0, 1, 1.0), (1, 2, 1.0), (2, 3, 1.0), (3, 0, 1.0), (0, 2, 1.4142135623730951),
             (1, 3, 1.4142135623730951), (0, 1, 9.0)]
    rows = [[(j, w) for i, j, w in edges if i == r] for r in range(4)]
    indptr = np.cumsum([0] + [len(r) for r in rows]).astype(np.int64)
    indices = np.array([j for r in rows for j, _ in r], np.int64)
    data = np.array([w for r in rows for _, w in r], np.float32)
    g = sp.csr_matrix((data, indices, indptr), shape=(4, 4))
    dgms, info = persistence(g, maxdim=1)
    assert canon(dgms) == SQ_L2 and info["method"] == "sparse"


# ---- the reference's Python parity cases, against its compiled persistence -------

@pytest.mark.parametrize("kw", [dict(maxdim=2), dict(maxdim=1, metric="l1"), dict(maxdim=1, metric="linf"),
                                dict(maxdim=1, metric="lp:1.5"), dict(maxdim=2, graph="knn", k=10),
                                dict(maxdim=1, threshold=0.7), dict(maxdim=0)])
def test_python_parity_cases(gpu, reference, kw):
    from turbotops_amd import persistence
    x = np.random.default_rng(0).normal(size=(600, 3)).astype(np.float32)
    print(f"case parity {kw}", flush=True)
    dgms, info = persistence(x, **kw)
    assert len(dgms) == kw["maxdim"] + 1
    res, _ = reference.persistence(x, maxdim=kw["maxdim"], threshold=kw.get("threshold", np.inf),
                                   metric=kw.get("metric", "l2"), graph=kw.get("graph", "full"), k=kw.get("k", 15))
    if kw.get("metric", "l2").startswith("lp:"):
        assert bars_close(dgms, ref_dgms(res, kw["maxdim"]), 1e-5)
    else:
        assert canon(dgms) == ref_canon(res)


def test_rejections(gpu):
    from turbotops_amd import persistence
    x = np.random.default_rng(0).normal(size=(50, 3)).astype(np.float32)
    for bad, msg in [(dict(metric="lp:0.5"), "p must be in"), (dict(graph="ring"), "graph must be"),
                     (dict(threshold=-1.0), "threshold must be non-negative"),
                     (dict(threshold=float("nan")), "threshold must be non-negative"),
                     (dict(graph="knn", k=0), "k must be >= 1"), (dict(maxdim=31), "maxdim must be <= 30"),
                     (dict(graph="knn", approximate=True), "without HNSW"),
                     (dict(memory_budget=500), "exceeds the memory budget")]:
        with pytest.raises(ValueError, match=msg):
            persistence(x, **bad)
    y = x.copy()
    y[4, 2] = np.inf
    with pytest.raises(ValueError, match="non-finite coordinate"):
        persistence(y)
    with pytest.raises(ValueError, match="2-d array"):
        persistence(x[0])


# ---- self-test consistency checks ----------------------------------------------

def torus(n, seed, R=2.0, r=1.0):
    rng = np.random.default_rng(seed)
    a, b = rng.uniform(0, 2 * np.pi, size=(2, n))
    return np.stack([(R + r * np.cos(a)) * np.cos(b), (R + r * np.cos(a)) * np.sin(b), r * np.sin(a)], 1).astype(
        np.float32)


def sphere(n, dim, seed):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(n, dim))
    return (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)


def test_mst_h0_equals_rips_h0_and_dense_equals_complete_graph(gpu):
    from turbotops_amd import persistence, radius_graph
    x = torus(160, 3)
    d2, info = persistence(x, maxdim=2)
    d0, _ = persistence(x, maxdim=0)
    assert canon(d2[:1]) == canon(d0)
    g = radius_graph(x, "l2", 1e30)
    import scipy.sparse as sp
    csr = sp.csr_matrix((g.weights, g.col_idx.astype(np.int64), g.row_ptr.astype(np.int64)), shape=(160, 160))
    dg, _ = persistence(csr, maxdim=2, threshold=info["threshold"])
    assert canon(dg) == canon(d2)


def test_sphere_has_one_dominant_h2(gpu):
    from turbotops_amd import persistence
    dgms, _ = persistence(sphere(120, 3, 5), maxdim=2)
    assert not np.isinf(dgms[1][:, 1]).any() and not np.isinf(dgms[2][:, 1]).any()
    assert int(((dgms[2][:, 1] - dgms[2][:, 0]) > 0.5).sum()) == 1


def test_lower_layout_equals_full(gpu):
    from turbotops_amd import persistence
    x = torus(120, 13)
    full, _ = persistence(x, maxdim=2)
    budget = (120 * 120 * 4 + 120 * 119 * 2) // 2
    low, _ = persistence(x, maxdim=2, memory_budget=budget)
    assert canon(low) == canon(full)


def _selftest_cloud(kind, n, dim):
    if kind == "grid":
        return np.random.default_rng(9).integers(0, 4, size=(n, dim)).astype(np.float32)
    rng = np.random.default_rng(21)
    if kind == "torus":
        return torus(n, 21)
    if kind == "sphere":
        return sphere(n, dim, 21)
    if kind == "uniform":
        return rng.uniform(size=(n, dim)).astype(np.float32)
    return rng.normal(size=(n, dim)).astype(np.float32)


@pytest.mark.parametrize("kind,n,dim,maxdim,metric", [
    ("torus", 150, 3, 2, "l2"), ("sphere", 70, 4, 3, "l2"), ("uniform", 40, 3, 4, "l2"), ("gauss", 90, 5, 2, "l1"),
    ("gauss", 90, 5, 2, "linf"), ("gauss", 90, 5, 2, "lp:3"), ("grid", 64, 3, 2, "l1"), ("grid", 64, 3, 3, "linf"),
    ("torus", 400, 3, 1, "l2"), ("uniform", 300, 3, 2, "l2"), ("grid", 200, 2, 2, "l2"), ("sphere", 150, 3, 2, "l1")])
def test_selftest_clouds_vs_reference(gpu, reference, kind, n, dim, maxdim, metric):
    """The twelve clouds of the reference's engine-parity check
    (src/turbotops.cpp:573-619): dense, kNN-8 and radius graph at 0.6 x the
    enclosing radius, against the compiled reference bar for bar."""
    from turbotops_amd import persistence
    x = _selftest_cloud(kind, n, dim)
    print(f"case selftest {kind} n={n} d={dim} maxdim={maxdim} {metric}", flush=True)
    lp = metric.startswith("lp:")
    dgms, info = persistence(x, maxdim=maxdim, metric=metric)
    res, _ = reference.persistence(x, maxdim=maxdim, metric=metric)
    assert bars_close(dgms, ref_dgms(res, maxdim), 1e-5) if lp else canon(dgms) == ref_canon(res)
    dk, _ = persistence(x, maxdim=maxdim, metric=metric, graph="knn", k=8)
    rk, _ = reference.persistence(x, maxdim=maxdim, metric=metric, graph="knn", k=8)
    assert bars_close(dk, ref_dgms(rk, maxdim), 1e-5) if lp else canon(dk) == ref_canon(rk)
    tau = float(np.float32(0.6) * np.float32(info["threshold"]))
    dr, _ = persistence(x, maxdim=maxdim, metric=metric, threshold=tau)
    rr, _ = reference.persistence(x, maxdim=maxdim, metric=metric, threshold=tau)
    assert bars_close(dr, ref_dgms(rr, maxdim), 1e-5) if lp else canon(dr) == ref_canon(rr)


def test_engine_choice(gpu):
    """maxdim 1 of a dense cloud runs the boundary engine (the reference's
    auto choice, include/turbotops/turbotops.hpp:144-152), maxdim 2 and
    mode-free sparse input the cohomology engine; H0 and H1 agree."""
    from turbotops_amd import persistence
    x = np.random.default_rng(4).normal(size=(500, 3)).astype(np.float32)
    d1, i1 = persistence(x, maxdim=1)
    d2, i2 = persistence(x, maxdim=2)
    assert i1["method"] == "dense-redzed" and i1["engine"] == "boundary"
    assert i2["method"] == "dense" and i2["engine"] == "coboundary"
    assert canon(d1) == canon(d2[:2])


def test_torus_c1_bar_for_bar(gpu, reference):
    """The plain CPU reference configuration: n = 1000 points on a torus in
    R^3 (R = 1, r = 0.3, angles U[0, 2pi)^2, noise N(0, 0.01^2), seed 1),
    full Rips, L2, maxdim 1."""
    from turbotops_amd import persistence
    rng = np.random.default_rng(1)
    th, ph = rng.uniform(0, 2 * np.pi, size=(2, 1000))
    x = np.stack([(1 + 0.3 * np.cos(th)) * np.cos(ph), (1 + 0.3 * np.cos(th)) * np.sin(ph), 0.3 * np.sin(th)], 1)
    x = (x + rng.normal(scale=0.01, size=x.shape)).astype(np.float32)
    dgms, info = persistence(x, maxdim=1)
    res, _ = reference.persistence(x, maxdim=1)
    assert canon(dgms) == ref_canon(res)


def _lp64(x, p):
    """Lp distances of every pair recomputed in float64 from the float32
    points (the exact formula, no intermediate float rounding), then rounded
    once to float32."""
    x64 = x.astype(np.float64)
    diff = np.abs(x64[:, None, :] - x64[None, :, :])
    return (np.power(np.power(diff, p).sum(-1), 1.0 / p)).astype(np.float32)


@pytest.mark.parametrize("p,n,dim,seed", [(1.5, 220, 4, 61), (3.3, 180, 6, 62), (1.5, 160, 16, 63)])
def test_lp_against_fp64_recomputation(gpu, reference, restatement, p, n, dim, seed):
    """Lp is matched within 2 ulp per distance, not bitwise, so its diagrams are
    also checked against an FP64 recomputation. The reference's float32
    algorithm itself (the oracle's restatement of it) lands within 3 ulp of
    the float64 distances rounded once on these clouds; the GPU's matrix must
    be within 4, and the GPU's diagram must equal, bar for bar within 6 ulp,
    the reference engine's diagram of the complete graph weighted by the
    float64 distances (the same filtration up to the enclosing radius, which
    changes no bar)."""
    from turbotops_amd import make_dense, persistence
    x = np.random.default_rng(seed).uniform(size=(n, dim)).astype(np.float32)
    metric = f"lp:{p}"
    print(f"case lp fp64 p={p} n={n} dim={dim}", flush=True)
    d64 = _lp64(x, p)
    _, mat, _ = make_dense(x, metric)
    off = ~np.eye(n, dtype=bool)
    assert ulp_diff(restatement.make_dense(x, metric)[off], d64[off]).max() <= 3
    assert ulp_diff(mat[off], d64[off]).max() <= 4
    dg, _ = persistence(x, maxdim=1, metric=metric)
    iu, ju = np.nonzero(off)
    rp = np.searchsorted(iu, np.arange(n + 1)).astype(np.uint64)
    (dims, b, e) = reference.persistence_csr(n, rp, ju.astype(np.uint32), d64[off], maxdim=1)
    want = sorted(zip(dims.tolist(), b.tolist(), e.tolist()))
    got = sorted(canon_tuples(dg))
    want = [t for t in want if t[1] != t[2]]
    assert len(got) == len(want)
    for (d1, b1, e1), (d2, b2, e2) in zip(got, want):
        assert d1 == d2 and math.isinf(e1) == math.isinf(e2)
        for u, v in ((b1, b2), (e1, e2)):
            if math.isfinite(v):
                assert ulp_diff(np.float32(u), np.float32(v)).max() <= 6, (d1, u, v)
