"""turbo_tops.persistence() on the MI355X front end.

Same signature, coercions, return value and errors as the reference's
Python entry point (python/turbo_tops/__init__.py:23-71): a point cloud
(n, d) under a Minkowski metric, or any object with .tocsr() (a sparse
weighted graph). Returns (dgms, info): dgms is a list of (m, 2) float32
arrays of (birth, death), one per dimension 0..maxdim (death = inf for
essential classes); info has method, threshold, edges, index_bits,
seconds_graph, seconds_total and peak_rss_mb.
"""
import math

import numpy as np

from . import capi


def _ext():
    capi.lib()   # libttgpu.so must be there (the module links it)
    from . import _turbotops
    return _turbotops


def persistence(x, maxdim=1, threshold=math.inf, metric="l2", graph="full", k=15,
                approximate=False, recall_samples=0, threads=0, memory_budget=0, *, _mode="auto",
                _force_wide=False):
    """Persistence diagrams of the Vietoris-Rips filtration.

    x: array (n, d) or scipy.sparse matrix (n, n); maxdim: highest homology
    dimension; threshold: filtration cutoff (inclusive; with graph="full"
    and threshold=inf the filtration stops at the enclosing radius, which
    changes no bars); metric: "l1", "l2", "linf" or "lp:<p>", p >= 1;
    graph: "full" or "knn"; k: neighbours for graph="knn"; approximate:
    HNSW kNN (not part of this build: raises); threads: host worker
    threads; memory_budget: bytes for the host copy of the dense matrix
    (0 = half of RAM). Not in the reference's signature, for tests: _mode
    ("auto", "boundary", "coboundary": the engine, as the reference CLI's
    --mode) and _force_wide (128-bit simplex indices in the cohomology
    engine, as rips::compute's force_wide)."""
    ext = _ext()
    if hasattr(x, "tocsr"):
        g = x.tocsr()
        g.sort_indices()
        return ext.persistence_csr(
            np.ascontiguousarray(g.indptr, dtype=np.int64),
            np.ascontiguousarray(g.indices, dtype=np.int64),
            np.ascontiguousarray(g.data, dtype=np.float32),
            int(g.shape[0]), int(maxdim), float(threshold), int(threads), bool(_force_wide))
    x = np.ascontiguousarray(x, dtype=np.float32)
    if x.ndim != 2:
        raise ValueError("x must be a 2-d array (n, d)")
    return ext.persistence_points(
        x, int(maxdim), float(threshold), str(metric), str(graph), int(k),
        bool(approximate), int(recall_samples), int(threads), int(memory_budget), str(_mode),
        bool(_force_wide))


def persistence_pairs(x, maxdim=1, threshold=math.inf, metric="l2", graph="full", k=15, threads=0,
                      memory_budget=0):
    """persistence() plus, per bar, the critical edges (i, j) of its birth
    and death simplices: the edge whose length is the simplex's filtration
    value ((-1, -1) where there is none). Returns (dgms, birth_edges,
    death_edges, info)."""
    ext = _ext()
    x = np.ascontiguousarray(x, dtype=np.float32)
    if x.ndim != 2:
        raise ValueError("x must be a 2-d array (n, d)")
    return ext.persistence_pairs(x, int(maxdim), float(threshold), str(metric), str(graph), int(k),
                                 int(threads), int(memory_budget))


class BatchPersistence:
    """Many clouds of uneven size and dimension through one batched GPU front
    end (dense full Rips at each cloud's enclosing radius), with host
    reductions in parallel. Create once per set of shapes; call with the
    clouds (a list of (n_i, d_i) arrays) for a list of (dgms, info), each equal
    to persistence(cloud, maxdim=maxdim, metric=metric). At maxdim 1, mode
    "auto" and "boundary" reduce with the boundary engine, "coboundary" with
    the cohomology engine (same diagrams); maxdim 0 and >= 2 run the
    cohomology engine (maxdim >= 2 on every kept edge with its GPU apparent
    flag). witness_flags (the default) adds the
    GPU's ready-witness flags to the boundary engine's inputs: one more device
    scan (about 0.12 s on 512 clouds) that lets the host reduction skip most
    edges of clouds without long cycles; witness_flags=False leaves it out.

    graph="knn": each cloud's flag complex of its symmetric exact k-NN graph
    instead (the kNN stage of the batch: every cloud's lists and graph in one
    device sequence; any k >= 1, lists longer than 64 through the
    single-cloud selection path, run without the recorded graph), each result
    equal to
    persistence(cloud, maxdim=maxdim, metric=metric, graph="knn", k=k);
    graph(c) and lists(c) give cloud c's graph and lists after a call."""

    def __init__(self, shapes, metric="l2", maxdim=1, mode="auto", witness_flags=True, graph="full", k=15):
        ext = _ext()
        self.shapes = [(int(n), int(d)) for n, d in shapes]
        ns, ds = [n for n, _ in self.shapes], [d for _, d in self.shapes]
        if graph not in ("full", "knn"):
            raise ValueError(f"graph must be 'full' or 'knn', not {graph!r}")
        self.graph_kind = graph
        if graph == "knn":
            self._r = ext.BatchKnnRunner(ns, ds, str(metric), int(k), int(maxdim), str(mode))
        else:
            self._r = ext.BatchRunner(ns, ds, str(metric), int(maxdim), str(mode), bool(witness_flags))

    def __call__(self, clouds, use_graph=True, threads=0, pairs=False):
        """With pairs=True each entry is (dgms, birth_edges, death_edges, info)
        as persistence_pairs returns it."""
        clouds = [np.ascontiguousarray(c, dtype=np.float32) for c in clouds]
        if [c.shape for c in clouds] != self.shapes:
            raise ValueError("clouds do not match the batch shapes")
        return self._r.run(clouds, bool(use_graph), int(threads), bool(pairs))

    def front_end(self, clouds, use_graph=True):
        """The batched device front end alone (no host reduction); the
        results land in host memory before it returns. See device_ms."""
        clouds = [np.ascontiguousarray(c, dtype=np.float32) for c in clouds]
        if [c.shape for c in clouds] != self.shapes:
            raise ValueError("clouds do not match the batch shapes")
        self._r.front(clouds, bool(use_graph))

    @property
    def device_ms(self):
        return self._r.device_ms

    def time_distance_stage(self):
        """The distance stage alone, re-run on the last points: the median of
        five launches after a warm-up (ms)."""
        return self._r.time_distance_stage()

    def graph(self, cloud):
        """graph="knn": cloud `cloud`'s symmetric kNN graph of the last call as
        a CsrGraph."""
        from .frontend import CsrGraph
        rp, col, w = self._r.graph(int(cloud))
        return CsrGraph(self.shapes[cloud][0], rp, col, w)

    def lists(self, cloud):
        """graph="knn": cloud `cloud`'s (n, min(k, n - 1)) neighbour lists."""
        return self._r.lists(int(cloud))

    @property
    def nnz(self):
        """graph="knn": CSR entries of every cloud's graph in the last call."""
        return self._r.nnz

    @property
    def edges(self):
        """(kept edges, kept edges not in an apparent pair) of the last run."""
        return self._r.edges_total, self._r.edges_kept

    @property
    def rows_per_wave(self):
        """Tuned rows per wave of the distance tiles (of the dimension group
        with the most tiles; each group has its own)."""
        return self._r.rows_per_wave


def persistence_batch(clouds, maxdim=1, metric="l2", threads=0, use_graph=True, graph="full", k=15):
    """persistence() of every cloud in `clouds` through one batched launch."""
    shapes = [np.shape(c) for c in clouds]
    bp = BatchPersistence(shapes, metric=metric, maxdim=maxdim, graph=graph, k=k)
    return bp(clouds, use_graph=use_graph, threads=threads)
