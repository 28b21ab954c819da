"""The batched front end (tt_batch_*, BatchPersistence): every cloud's
diagram equals persistence() on that cloud alone (and so the reference's),
with and without the HIP graph, over repeated replays with eager work in
between."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def canon(dgms):
    bars = sorted((d, float(b), float(e)) for d, a in enumerate(dgms) for b, e in np.asarray(a))
    return [f"{d} {b:.9g} inf" if math.isinf(e) else f"{d} {b:.9g} {e:.9g}" for d, b, e in bars]


def _clouds(seed, count=24, nmax=400):
    from turbotops_amd.synthetic import uneven_clouds
    cl = uneven_clouds(count=count, nmin=2, nmax=nmax, seed=seed)
    cl[0] = cl[0][:1]          # a single point
    cl[1] = cl[1][:2]          # a single edge
    return cl


@pytest.mark.parametrize("metric", ["l2", "linf"])
def test_batch_equals_single(gpu, metric):
    from turbotops_amd import BatchPersistence, persistence
    clouds = _clouds(11)
    bp = BatchPersistence([c.shape for c in clouds], metric=metric)
    print(f"case batch metric={metric} clouds={len(clouds)}", flush=True)
    for use_graph in (False, True):
        res = bp(clouds, use_graph=use_graph)
        assert len(res) == len(clouds)
        for c, (dgms, info) in zip(clouds, res):
            want, _ = persistence(c, maxdim=1, metric=metric)
            assert canon(dgms) == canon(want)
    assert bp.rows_per_wave in (8, 16, 32)


def test_batch_engines_agree(gpu):
    """The boundary engine (every edge in (diameter, comb index) order, with
    the GPU's ready-witness flags by default, and without them) and the cohomology
    engine (the non-apparent edges in its own order) on the same batch: the
    same diagrams, with and without the graph, and switching back and forth
    between handles."""
    from turbotops_amd import BatchPersistence
    clouds = _clouds(21, count=20, nmax=500)
    print(f"case batch engines clouds={len(clouds)}", flush=True)
    bd = BatchPersistence([c.shape for c in clouds], witness_flags=False)
    bw = BatchPersistence([c.shape for c in clouds])
    bc = BatchPersistence([c.shape for c in clouds], mode="coboundary")
    for use_graph in (True, False, True):
        rd = bd(clouds, use_graph=use_graph)
        rw = bw(clouds, use_graph=use_graph)
        rc = bc(clouds, use_graph=use_graph)
        for (dd, id_), (dw, iw), (dc, ic) in zip(rd, rw, rc):
            assert canon(dd) == canon(dc) == canon(dw)
            assert id_["engine"] == iw["engine"] == "boundary" and ic["engine"] == "coboundary"


def test_graph_replays_with_new_points(gpu):
    """Replays of the recorded graph on new data of the same shapes, with
    eager library work in between, each checked against the eager result."""
    from turbotops_amd import BatchPersistence, make_dense, persistence
    base = _clouds(12, count=16, nmax=300)
    bp = BatchPersistence([c.shape for c in base], metric="l2")
    for rep in range(4):
        rng = np.random.default_rng(100 + rep)
        clouds = [(c + rng.normal(scale=0.05, size=c.shape)).astype(np.float32) for c in base]
        res = bp(clouds, use_graph=True)
        make_dense(clouds[3], "l2")        # eager work between replays
        for c, (dgms, _) in zip(clouds, res):
            want, _ = persistence(c, maxdim=1)
            assert canon(dgms) == canon(want)


def test_batch_vs_reference(gpu, reference):
    from turbotops_amd import BatchPersistence
    clouds = _clouds(13, count=8, nmax=250)
    res = BatchPersistence([c.shape for c in clouds])(clouds)
    for c, (dgms, _) in zip(clouds, res):
        dims, b, e = reference.persistence(c, maxdim=1)[0]
        want = [np.stack([b[dims == d], e[dims == d]], 1) for d in range(2)]
        assert canon(dgms) == canon(want)


@pytest.mark.parametrize("maxdim", [0, 2, 3])
def test_batch_other_maxdim(gpu, maxdim):
    """maxdim 0 (union-find on the non-apparent edges) and maxdim 2, 3 (every
    kept edge with its GPU apparent flag into the cohomology engine, order 3):
    each cloud's diagrams equal persistence() on that cloud alone, with and
    without the graph."""
    from turbotops_amd import BatchPersistence, persistence
    clouds = _clouds(15, count=12, nmax=160 if maxdim == 3 else 300)
    print(f"case batch maxdim={maxdim} clouds={len(clouds)}", flush=True)
    bp = BatchPersistence([c.shape for c in clouds], maxdim=maxdim)
    for use_graph in (True, False):
        res = bp(clouds, use_graph=use_graph)
        for c, (dgms, info) in zip(clouds, res):
            want, _ = persistence(c, maxdim=maxdim)
            assert len(dgms) == maxdim + 1
            assert canon(dgms) == canon(want)
            assert info["engine"] == "coboundary"


def test_batch_maxdim2_vs_reference(gpu, reference):
    from turbotops_amd import BatchPersistence
    clouds = _clouds(16, count=6, nmax=200)
    res = BatchPersistence([c.shape for c in clouds], maxdim=2)(clouds)
    for c, (dgms, _) in zip(clouds, res):
        dims, b, e = reference.persistence(c, maxdim=2)[0]
        want = [np.stack([b[dims == d], e[dims == d]], 1) for d in range(3)]
        assert canon(dgms) == canon(want)


def test_batch_rejects(gpu):
    from turbotops_amd import BatchPersistence
    clouds = _clouds(14, count=4, nmax=50)
    with pytest.raises(ValueError, match="maxdim must be <= 30"):
        BatchPersistence([c.shape for c in clouds], maxdim=31)
    bp = BatchPersistence([c.shape for c in clouds])
    with pytest.raises(ValueError, match="do not match"):
        bp(clouds[:3])
    bad = [c.copy() for c in clouds]
    bad[2][0, 0] = np.nan
    with pytest.raises(ValueError, match="non-finite"):
        bp(bad)
