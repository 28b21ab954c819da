"""The cohomology engine's host reduction (turbotops_amd/host/cohomology.hpp)
on the CPU, fed by the oracle: dense full matrix, edges in filtration order
and apparent flags from the restatement; sparse graphs from the
restatement's radius and kNN graphs; the columns of every dimension above 1
and all emergent states from the restatement of the GPU's column stages
(tto_assemble_cx, tto_emergent_cx). Diagrams must equal the reference's (the
CLI goldens, and the compiled reference's tops::persistence) bar for bar,
with 64-bit indices and with 128-bit ones forced (the reference's own
force_wide self-test, src/turbotops.cpp:446-450)."""
import ctypes as C
import math
import os

import numpy as np
import pytest

from conftest import ROOT

_f32p = C.POINTER(C.c_float)
_u32p = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)
_u8p = C.POINTER(C.c_uint8)


@pytest.fixture(scope="module")
def harness():
    path = os.path.join(ROOT, "tests", "libhost_harness.so")
    if not os.path.exists(path):
        pytest.skip("tests/libhost_harness.so not built (make -C tests)")
    L = C.CDLL(path)
    out = [_u64p, C.POINTER(_u32p), C.POINTER(_f32p), C.POINTER(_f32p)]
    L.hh_dense.argtypes = [_f32p, C.c_uint32, C.c_float, C.c_uint32, _u64p, _f32p, _u8p, C.c_uint64,
                           C.c_int] + out
    L.hh_sparse.argtypes = [C.c_uint32, _u64p, _u32p, _f32p, C.c_float, C.c_uint32, _u64p, _f32p, _u8p,
                            C.c_uint64, C.c_int] + out
    L.hh_last_error.restype = C.c_char_p
    L.hh_free.argtypes = [C.c_void_p]
    return L


def _p(a, t):
    return a.ctypes.data_as(t)


def _take(L, nb, dims, b, e):
    k = nb.value
    out = []
    for p, dt in ((dims, np.uint32), (b, np.float32), (e, np.float32)):
        out.append(np.ctypeslib.as_array(p, shape=(max(k, 1),))[:k].astype(dt, copy=True))
        L.hh_free(C.cast(p, C.c_void_p))
    return out


def canon(dims, b, e):
    bars = sorted(zip(dims.tolist(), b.tolist(), e.tolist()))
    return [f"{d} {x:.9g} inf" if math.isinf(y) else f"{d} {x:.9g} {y:.9g}" for d, x, y in bars]


def dense_diagram(L, R, x, metric, maxdim, wide=False):
    mat = R.make_dense(x, metric, True)
    n = x.shape[0]
    thr = R.enclosing_radius(mat, n)
    idx, d = R.dense_edges(mat, thr)
    ap = R.apparent_edges(mat, idx, d)
    nb, dims, b, e = C.c_uint64(), _u32p(), _f32p(), _f32p()
    rc = L.hh_dense(_p(mat, _f32p), n, thr, maxdim, _p(idx, _u64p), _p(d, _f32p), _p(ap, _u8p), idx.size,
                    int(wide), C.byref(nb), C.byref(dims), C.byref(b), C.byref(e))
    assert rc == 0, L.hh_last_error()
    return _take(L, nb, dims, b, e)


def sparse_diagram(L, n, rp, col, w, maxdim, threshold=np.inf, wide=False):
    rp = np.ascontiguousarray(rp, np.uint64)
    col = np.ascontiguousarray(col, np.uint32)
    w = np.ascontiguousarray(w, np.float32)
    edges = []
    for v in range(n):
        for t in range(int(rp[v]), int(rp[v + 1])):
            u = int(col[t])
            if u < v and w[t] <= threshold:
                edges.append((float(w[t]), -(v * (v - 1) // 2 + u)))
    edges.sort()
    idx = np.array([-e[1] for e in edges], np.uint64)
    d = np.array([e[0] for e in edges], np.float32)
    from oracle import Complex, Restatement
    cx = Complex(n=n, row_ptr=rp, col=col, w=w)
    ap = np.zeros(max(idx.size, 1), np.uint8)
    if idx.size:
        ap[: idx.size] = Restatement().apparent_cx(cx, 2, [int(v) for v in idx], d) & 1
    nb, dims, b, e = C.c_uint64(), _u32p(), _f32p(), _f32p()
    rc = L.hh_sparse(n, _p(rp, _u64p), _p(col, _u32p), _p(w, _f32p), threshold, maxdim, _p(idx, _u64p),
                     _p(d, _f32p), _p(ap, _u8p), idx.size, int(wide), C.byref(nb), C.byref(dims), C.byref(b),
                     C.byref(e))
    assert rc == 0, L.hh_last_error()
    return _take(L, nb, dims, b, e)


SQUARE = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32)


@pytest.mark.parametrize("metric,gold", [
    ("l2", ["0 0 1", "0 0 1", "0 0 1", "0 0 inf", "1 1 1.41421354"]),
    ("l1", ["0 0 1", "0 0 1", "0 0 1", "0 0 inf", "1 1 2"]),
    ("linf", ["0 0 1", "0 0 1", "0 0 1", "0 0 inf"]),
    ("lp:3", ["0 0 1", "0 0 1", "0 0 1", "0 0 inf", "1 1 1.25992107"])])
def test_square_goldens(harness, restatement, metric, gold):
    # tests/test_cli.sh:56-94
    assert canon(*dense_diagram(harness, restatement, SQUARE, metric, 1)) == gold
    assert canon(*dense_diagram(harness, restatement, SQUARE, metric, 2)) == gold


def test_square_threshold_golden(harness, restatement):
    # sq.thr.gold: the radius graph at tau = 1 keeps the 4-cycle, an essential H1
    rp, col, w = restatement.radius_graph(SQUARE, "l2", 1.0)
    got = canon(*sparse_diagram(harness, 4, rp, col, w, 1, 1.0))
    assert got == ["0 0 1", "0 0 1", "0 0 1", "0 0 inf", "1 1 inf"]


def _cloud(kind, n, dim, seed):
    rng = np.random.default_rng(seed)
    if kind == "grid":
        return rng.integers(0, 4, size=(n, dim)).astype(np.float32)
    if kind == "torus":
        a, b = rng.uniform(0, 2 * np.pi, size=(2, n))
        return np.stack([(2 + np.cos(a)) * np.cos(b), (2 + np.cos(a)) * np.sin(b), np.sin(a)], 1).astype(np.float32)
    if kind == "sphere":
        v = rng.normal(size=(n, dim))
        return (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)
    if kind == "uniform":
        return rng.uniform(size=(n, dim)).astype(np.float32)
    return rng.normal(size=(n, dim)).astype(np.float32)


CASES = [("torus", 90, 3, 2, "l2"), ("sphere", 50, 4, 3, "l2"), ("uniform", 30, 3, 4, "l2"),
         ("gauss", 60, 5, 2, "l1"), ("gauss", 60, 5, 2, "linf"), ("grid", 48, 3, 2, "l1"),
         ("grid", 40, 3, 3, "linf"), ("torus", 200, 3, 1, "l2"), ("grid", 100, 2, 2, "l2"),
         ("sphere", 80, 3, 2, "l1"), ("gauss", 120, 3, 1, "lp:3")]


@pytest.mark.parametrize("kind,n,dim,maxdim,metric", CASES)
def test_dense_vs_reference(harness, restatement, reference, kind, n, dim, maxdim, metric):
    x = _cloud(kind, n, dim, n + dim)
    got = canon(*dense_diagram(harness, restatement, x, metric, maxdim))
    res, _ = reference.persistence(x, maxdim=maxdim, metric=metric, mode="coboundary")
    assert got == canon(*res)


@pytest.mark.parametrize("kind,n,dim,maxdim,metric", CASES)
def test_sparse_vs_reference(harness, restatement, reference, kind, n, dim, maxdim, metric):
    x = _cloud(kind, n, dim, n + dim)
    ids = restatement.knn(x, metric, 8)
    rp, col, w = restatement.knn_graph(x, metric, ids)
    got = canon(*sparse_diagram(harness, n, rp, col, w, maxdim))
    res, _ = reference.persistence(x, maxdim=maxdim, metric=metric, graph="knn", k=8, mode="coboundary")
    assert got == canon(*res)
    mat = restatement.make_dense(x, metric)
    tau = float(np.float32(0.6) * restatement.enclosing_radius(mat, n))
    rp, col, w = restatement.radius_graph(x, metric, tau)
    got = canon(*sparse_diagram(harness, n, rp, col, w, maxdim))
    res, _ = reference.persistence(x, maxdim=maxdim, metric=metric, threshold=tau, mode="coboundary")
    assert got == canon(*res)


def test_sparse_knn_graph_larger(harness, restatement, reference):
    """A kNN-12 graph of 3000 points in [0,1]^8 at maxdim 2: many columns
    per dimension, so later columns keep meeting earlier pivots (the stored
    reduced coboundaries) and the working columns run on the radix heap; bar
    for bar against the compiled reference."""
    x = np.random.default_rng(31).uniform(size=(3000, 8)).astype(np.float32)
    rp, col, w = reference.knn_graph(x, "l2", 12)
    got = canon(*sparse_diagram(harness, 3000, rp, col, w, 2))
    res, _ = reference.persistence(x, maxdim=2, metric="l2", graph="knn", k=12, mode="coboundary")
    assert got == canon(*res)
    assert sum(1 for s in got if s.startswith("1 ")) > 1000


@pytest.mark.parametrize("kind,n,dim,maxdim,metric", [c for c in CASES if c[3] >= 2] + [("torus", 90, 3, 1, "l2")])
def test_forced_128_bit_indices_equal_64_bit(harness, restatement, reference, kind, n, dim, maxdim, metric):
    """128-bit simplex indices forced (the u128 engine with the binary heap,
    which 64-bit indices never reach here) give the 64-bit diagrams, dense and
    sparse, as the reference's self-test requires of its own engine
    (src/turbotops.cpp:446-450); the reference's forced-wide compute agrees."""
    from oracle import Complex
    x = _cloud(kind, n, dim, n + dim)
    narrow = canon(*dense_diagram(harness, restatement, x, metric, maxdim))
    assert canon(*dense_diagram(harness, restatement, x, metric, maxdim, wide=True)) == narrow
    mat = restatement.make_dense(x, metric)
    (bars, wide) = reference.compute_cx(Complex(mat=mat), restatement.enclosing_radius(mat, n), maxdim, True)
    assert wide and canon(*bars) == narrow
    rp, col, w = restatement.knn_graph(x, metric, restatement.knn(x, metric, 8))
    s64 = canon(*sparse_diagram(harness, n, rp, col, w, maxdim))
    assert canon(*sparse_diagram(harness, n, rp, col, w, maxdim, wide=True)) == s64
