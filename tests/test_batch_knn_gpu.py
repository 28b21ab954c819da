"""The kNN stage of the batch (tt_batch_knn_*, BatchPersistence(graph="knn")):
every cloud's lists and symmetric graph equal the single-cloud device path
(tt_knn_exact / tt_knn_graph_exact, themselves pinned against the oracle's
brute force and the reference in test_graph_gpu.py) bit for bit, for every
metric kind, specialised and runtime dimensions, and clouds smaller than
k + 1; every cloud's diagrams equal persistence(cloud, graph="knn", k) and,
on a sample, the compiled reference's; the recorded graph replays on new
points."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def canon(dgms):
    bars = sorted((d, float(b), float(e)) for d, a in enumerate(dgms) for b, e in np.asarray(a))
    return [f"{d} {b:.9g} inf" if math.isinf(e) else f"{d} {b:.9g} {e:.9g}" for d, b, e in bars]


def _clouds(seed, count=16, nmax=600, dims=(2, 3, 8, 5)):
    rng = np.random.default_rng(seed)
    out = []
    for t in range(count):
        n = int(rng.integers(2, nmax))
        d = int(dims[t % len(dims)])
        out.append(rng.uniform(size=(n, d)).astype(np.float32))
    out[0] = out[0][:1]     # no neighbours
    out[1] = out[1][:2]     # k_c = 1
    out[2] = out[2][:9]     # k_c = 8 < k
    # ties: points on a small integer grid
    out[3] = rng.integers(0, 3, size=(200, out[3].shape[1])).astype(np.float32)
    return out


@pytest.mark.parametrize("metric,k", [("l2", 15), ("l1", 8), ("linf", 33), ("lp:1.5", 5), ("lp:3", 64),
                                      ("l2", 100), ("linf", 65), ("l1", 300)])
def test_graphs_equal_single_cloud(gpu, metric, k):
    """k above 64: clouds with more than k + 1 points take the selection path,
    smaller ones (k_c = n_c - 1 <= 64) the register-list scan, in one batch."""
    from turbotops_amd import BatchPersistence, knn_graph_exact, knn_lists_exact
    clouds = _clouds(31, count=14, nmax=500)
    print(f"case batch knn metric={metric} k={k} clouds={len(clouds)}", flush=True)
    bp = BatchPersistence([c.shape for c in clouds], metric=metric, graph="knn", k=k)
    bp.front_end(clouds)
    total = 0
    for i, c in enumerate(clouds):
        g = bp.graph(i)
        w = knn_graph_exact(c, metric, k)
        assert g.valid()
        assert np.array_equal(g.row_ptr, w.row_ptr), i
        assert np.array_equal(g.col_idx, w.col_idx), i
        assert np.array_equal(g.weights.view(np.uint32), w.weights.view(np.uint32)), i
        assert np.array_equal(bp.lists(i), knn_lists_exact(c, metric, k)), i
        total += g.col_idx.size
    assert bp.nnz == total


@pytest.mark.parametrize("maxdim", [1, 2])
def test_diagrams_equal_single_cloud(gpu, maxdim):
    from turbotops_amd import BatchPersistence, persistence
    clouds = _clouds(32, count=12, nmax=800)
    print(f"case batch knn diagrams maxdim={maxdim}", flush=True)
    bp = BatchPersistence([c.shape for c in clouds], graph="knn", k=10, maxdim=maxdim)
    for use_graph in (True, False):
        res = bp(clouds, use_graph=use_graph)
        for c, (dgms, info) in zip(clouds, res):
            want, _ = persistence(c, maxdim=maxdim, graph="knn", k=10)
            assert len(dgms) == maxdim + 1
            assert canon(dgms) == canon(want)


def test_diagrams_vs_reference(gpu, reference):
    from turbotops_amd import BatchPersistence
    clouds = _clouds(33, count=6, nmax=400, dims=(2, 3, 8))
    res = BatchPersistence([c.shape for c in clouds], graph="knn", k=12, maxdim=2)(clouds)
    for c, (dgms, _) in zip(clouds, res):
        dims, b, e = reference.persistence(c, maxdim=2, graph="knn", k=12)[0]
        want = [np.stack([b[dims == d], e[dims == d]], 1) for d in range(3)]
        assert canon(dgms) == canon(want)


def test_large_k_diagrams_and_reruns(gpu):
    """A batch with lists longer than 64 runs its sequence directly (use_graph
    is ignored) and gives the single-cloud diagrams on every run."""
    from turbotops_amd import BatchPersistence, persistence
    clouds = _clouds(35, count=8, nmax=400)
    print(f"case batch knn large k clouds={len(clouds)}", flush=True)
    bp = BatchPersistence([c.shape for c in clouds], graph="knn", k=90, maxdim=1)
    for rep in range(2):
        res = bp(clouds, use_graph=True)
        for c, (dgms, _) in zip(clouds, res):
            want, _ = persistence(c, maxdim=1, graph="knn", k=90)
            assert canon(dgms) == canon(want)


def test_replays_with_new_points(gpu):
    """Replays of the recorded graph on new data of the same shapes, with eager
    library work in between, each checked against the single-cloud graphs."""
    from turbotops_amd import BatchPersistence, knn_graph_exact
    base = _clouds(34, count=10, nmax=300)
    bp = BatchPersistence([c.shape for c in base], graph="knn", k=7)
    for rep in range(3):
        rng = np.random.default_rng(200 + rep)
        clouds = [(c + rng.normal(scale=0.05, size=c.shape)).astype(np.float32) for c in base]
        bp.front_end(clouds, use_graph=True)
        knn_graph_exact(clouds[5], "l2", 3)       # eager work between replays
        for i, c in enumerate(clouds):
            g, w = bp.graph(i), knn_graph_exact(c, "l2", 7)
            assert np.array_equal(g.col_idx, w.col_idx) and np.array_equal(g.row_ptr, w.row_ptr)
            assert np.array_equal(g.weights.view(np.uint32), w.weights.view(np.uint32))


def test_rejects(gpu):
    from turbotops_amd import BatchPersistence
    shapes = [(10, 3), (20, 2)]
    with pytest.raises(ValueError, match="k must be >= 1"):
        BatchPersistence(shapes, graph="knn", k=0)
    with pytest.raises(ValueError, match="graph must be"):
        BatchPersistence(shapes, graph="radius")
    bp = BatchPersistence(shapes, graph="knn", k=4)
    bad = [np.zeros((10, 3), np.float32), np.full((20, 2), np.nan, np.float32)]
    with pytest.raises(ValueError, match="non-finite"):
        bp(bad)
