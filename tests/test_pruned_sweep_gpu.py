"""The searches that follow a spatial order and skip by boxes (Boruvka's
nearest-outside-component search, the seeded kNN scan, the radius tiles)
against the oracle over seeded random clouds: shape (Gaussian, tight
clusters, integer grid with many ties, a segment, anisotropic), size, dimension
(runtime dimensions included), metric, k and tau drawn per seed. Bar: bit-exact
deaths, ids and CSR graphs (L1, L2, L-inf)."""
import numpy as np
import pytest

from conftest import bits

pytestmark = pytest.mark.gpu

SHAPES = ["gauss", "clusters", "grid", "segment", "aniso"]


def _cloud(rng, shape, n, d):
    if shape == "gauss":
        return rng.normal(size=(n, d)).astype(np.float32)
    if shape == "clusters":
        c = rng.uniform(-20, 20, size=(int(rng.integers(2, 9)), d))
        return (c[rng.integers(0, len(c), size=n)] + rng.normal(scale=0.05, size=(n, d))).astype(np.float32)
    if shape == "grid":
        return rng.integers(0, 5, size=(n, d)).astype(np.float32)
    if shape == "segment":
        t = rng.uniform(size=(n, 1))
        return (t * rng.normal(size=(1, d)) + rng.normal(size=(1, d))).astype(np.float32)
    return (rng.uniform(size=(n, d)) * rng.uniform(0.01, 50, size=(1, d))).astype(np.float32)


@pytest.mark.parametrize("seed", range(12))
def test_pruned_searches_sweep(gpu, restatement, seed):
    from turbotops_amd import knn_lists_exact, mst_deaths, radius_graph
    rng = np.random.default_rng(1000 + seed)
    shape = SHAPES[seed % len(SHAPES)]
    n = int(rng.integers(300, 5000))
    d = [1, 2, 3, 4, 5, 8][seed % 6]
    metric = str(rng.choice(["l1", "l2", "linf"]))
    x = _cloud(rng, shape, n, d)
    k = int(rng.integers(1, 65))
    print(f"case sweep seed={seed} shape={shape} n={n} d={d} metric={metric} k={k}", flush=True)
    assert np.array_equal(bits(mst_deaths(x, metric)), bits(restatement.mst_deaths(x, metric)))
    assert np.array_equal(knn_lists_exact(x, metric, k), restatement.knn(x, metric, k))
    sub = x[rng.choice(n, size=min(n, 400), replace=False)]
    tau = float(np.quantile(restatement.make_dense(sub, metric, True), rng.uniform(0.001, 0.05)))
    g = radius_graph(x, metric, tau)
    rp, col, w = restatement.radius_graph(x, metric, tau)
    assert np.array_equal(g.row_ptr, rp) and np.array_equal(g.col_idx, col)
    assert np.array_equal(bits(g.weights), bits(w))
