"""The boundary engine (turbotops_amd/host/zerodiff.hpp: H0 and H1 by
reduction to zero differentials, the reference's auto choice for H1 of dense
input, include/turbotops/redzed.hpp:1001-1411) on the CPU, fed by the oracle:
the restatement's full matrix and every edge up to the enclosing radius in
(diameter, combinatorial index) order. Pinned three ways:

* the reference CLI's square goldens (tests/test_cli.sh:56-94);
* brute force: the textbook reduction of the full boundary matrix (every
  edge and triangle, columns added left to right over Z/2) on clouds small
  enough to list all triangles;
* the compiled reference's diagrams, in both of its modes, bar for bar.

The engine's second input form, the ready-witness flag per edge (what
tt_build_edges_witness computes on the GPU), is checked the same ways, with
the flags from their definition; a planted mistake in the flags is seen.
"""
import ctypes as C
import itertools
import math
import os

import numpy as np
import pytest

from conftest import ROOT

_f32p = C.POINTER(C.c_float)
_u32p = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)


@pytest.fixture(scope="module")
def harness():
    path = os.path.join(ROOT, "tests", "libhost_harness.so")
    if not os.path.exists(path):
        pytest.skip("tests/libhost_harness.so not built (make -C tests)")
    L = C.CDLL(path)
    L.hh_boundary.argtypes = [_f32p, C.c_uint32, _u32p, _f32p, C.c_uint64, C.c_int, _u64p, C.POINTER(_u32p),
                              C.POINTER(_f32p), C.POINTER(_f32p)]
    L.hh_last_error.restype = C.c_char_p
    L.hh_free.argtypes = [C.c_void_p]
    return L


def comb_edges(mat, thr):
    """Every edge lo < hi with d <= thr, in (diameter, comb index) order."""
    n = mat.shape[0]
    hi, lo = np.tril_indices(n, -1)
    d = mat[hi, lo]
    keep = d <= thr
    hi, lo, d = hi[keep], lo[keep], d[keep]
    o = np.lexsort((hi.astype(np.int64) * (hi - 1) // 2 + lo, d))
    return np.stack([lo[o], hi[o]], 1).astype(np.uint32), np.ascontiguousarray(d[o], np.float32)


def witness_flags(mat, verts, d):
    """The ready-witness flag of every edge (lo, hi) of diameter D, from its
    definition: some w with d(lo, w) <= D and d(hi, w) <= D for which the
    edge is the youngest face of {lo, hi, w} in (diameter, comb index) order:
    w < lo; or lo < w < hi and d(hi, w) < D; or w > hi and both < D."""
    n = mat.shape[0]
    w = np.arange(n)[None, :]
    out = np.zeros(len(d), bool)
    for s0 in range(0, len(d), 1024):
        lo = verts[s0:s0 + 1024, 0].astype(np.int64)[:, None]
        hi = verts[s0:s0 + 1024, 1].astype(np.int64)[:, None]
        D = d[s0:s0 + 1024, None]
        ra, rb = mat[lo[:, 0]], mat[hi[:, 0]]
        both = (ra <= D) & (rb <= D) & (w != lo) & (w != hi)
        ready = (w < lo) | ((w > lo) & (w < hi) & (rb < D)) | ((w > hi) & (ra < D) & (rb < D))
        out[s0:s0 + 1024] = (both & ready).any(1)
    return out


def with_flags(verts, flags):
    v = np.array(verts, np.uint32)
    v[:, 1] |= np.where(flags, np.uint32(0x80000000), np.uint32(0))
    return v


def engine_bars(L, mat, verts, d, witness=False):
    mat = np.ascontiguousarray(mat, np.float32)
    verts = np.ascontiguousarray(verts, np.uint32)
    n = mat.shape[0]
    nb, dims, b, e = C.c_uint64(), _u32p(), _f32p(), _f32p()
    rc = L.hh_boundary(mat.ctypes.data_as(_f32p), n, verts.ctypes.data_as(_u32p), d.ctypes.data_as(_f32p),
                       d.size, int(witness), C.byref(nb), C.byref(dims), C.byref(b), C.byref(e))
    assert rc == 0, L.hh_last_error()
    k = nb.value
    out = []
    for p, dt in ((dims, np.uint32), (b, np.float32), (e, np.float32)):
        out.append(np.ctypeslib.as_array(p, shape=(max(k, 1),))[:k].astype(dt, copy=True))
        L.hh_free(C.cast(p, C.c_void_p))
    return out


def canon(dims, b, e):
    bars = sorted(zip(np.asarray(dims).tolist(), np.asarray(b).tolist(), np.asarray(e).tolist()))
    return [f"{d} {x:.9g} inf" if math.isinf(y) else f"{d} {x:.9g} {y:.9g}" for d, x, y in bars]


def engine_diagram(L, restatement, x, metric, witness=False):
    mat = restatement.make_dense(x, metric)
    thr = restatement.enclosing_radius(mat, x.shape[0]) if x.shape[0] else 0.0
    v, d = comb_edges(mat, thr)
    if witness:
        return engine_bars(L, mat, with_flags(v, witness_flags(mat, v, d)), d, True), mat, thr
    return engine_bars(L, mat, v, d), mat, thr


def brute_force(mat, thr):
    """H0 and H1 bars of the Rips filtration up to thr by the standard
    algorithm: simplices ordered by (diameter, dimension, vertex tuple),
    boundary columns over Z/2 reduced left to right until every pivot
    (lowest nonzero row) is unique; a pivot pair (i, j) is a bar
    (diam_i, diam_j) when the two differ, an unpaired creator is an
    essential bar. Triangles of diameter <= thr are included, so H1 classes
    born at or below thr die where the filtration kills them."""
    n = mat.shape[0]
    simp = [((0.0, 0, (v,))) for v in range(n)]
    for a, b in itertools.combinations(range(n), 2):
        if mat[a, b] <= thr:
            simp.append((float(mat[a, b]), 1, (a, b)))
    for a, b, c in itertools.combinations(range(n), 3):
        dt = max(mat[a, b], mat[a, c], mat[b, c])
        if dt <= thr:
            simp.append((float(dt), 2, (a, b, c)))
    simp.sort()
    pos = {s[2]: i for i, s in enumerate(simp)}
    cols = []
    for dg, dim, vs in simp:
        if dim == 0:
            cols.append(set())
        else:
            cols.append({pos[f] for f in itertools.combinations(vs, dim)})
    low_of = {}
    paired = set()
    bars = []
    for j, col in enumerate(cols):
        while col:
            low = max(col)
            if low in low_of:
                col ^= cols[low_of[low]]
            else:
                low_of[low] = j
                paired.update((low, j))
                bi, bj = simp[low][0], simp[j][0]
                if bi != bj:
                    bars.append((simp[low][1], bi, bj))
                break
    for i, (dg, dim, vs) in enumerate(simp):
        if i not in paired and not cols[i] and dim <= 1:
            bars.append((dim, dg, math.inf))
    dims, b, e = zip(*bars) if bars else ((), (), ())
    return np.array(dims), np.array(b, np.float32), np.array(e, np.float32)


SQUARE = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32)


@pytest.mark.parametrize("metric,gold", [
    ("l2", ["0 0 1", "0 0 1", "0 0 1", "0 0 inf", "1 1 1.41421354"]),
    ("l1", ["0 0 1", "0 0 1", "0 0 1", "0 0 inf", "1 1 2"]),
    ("linf", ["0 0 1", "0 0 1", "0 0 1", "0 0 inf"]),
    ("lp:3", ["0 0 1", "0 0 1", "0 0 1", "0 0 inf", "1 1 1.25992107"])])
def test_square_goldens(harness, restatement, metric, gold):
    # tests/test_cli.sh:56-94
    bars, _, _ = engine_diagram(harness, restatement, SQUARE, metric)
    assert canon(*bars) == gold


def _cloud(kind, n, dim, seed):
    rng = np.random.default_rng(seed)
    if kind == "grid":
        return rng.integers(0, 3, size=(n, dim)).astype(np.float32)
    if kind == "circle":
        t = rng.uniform(0, 2 * np.pi, n)
        x = np.zeros((n, dim))
        x[:, 0], x[:, 1] = np.cos(t), np.sin(t)
        return (x + rng.normal(scale=0.05, size=x.shape)).astype(np.float32)
    if kind == "torus":
        a, b = rng.uniform(0, 2 * np.pi, size=(2, n))
        return np.stack([(2 + np.cos(a)) * np.cos(b), (2 + np.cos(a)) * np.sin(b), np.sin(a)], 1).astype(np.float32)
    if kind == "uniform":
        return rng.uniform(size=(n, dim)).astype(np.float32)
    return rng.normal(size=(n, dim)).astype(np.float32)


@pytest.mark.parametrize("kind,n,dim,metric", [
    ("circle", 12, 2, "l2"), ("gauss", 14, 3, "l2"), ("grid", 13, 2, "l1"), ("grid", 14, 3, "linf"),
    ("uniform", 15, 4, "lp:1.5"), ("grid", 16, 2, "l2"), ("circle", 16, 3, "l1"), ("gauss", 9, 2, "linf")])
def test_brute_force(harness, restatement, kind, n, dim, metric):
    x = _cloud(kind, n, dim, 7 * n + dim)
    bars, mat, thr = engine_diagram(harness, restatement, x, metric)
    assert canon(*bars) == canon(*brute_force(mat, thr))


CASES = [("circle", 400, 2, "l2"), ("circle", 300, 3, "l1"), ("torus", 300, 3, "l2"), ("gauss", 250, 5, "linf"),
         ("grid", 125, 3, "l2"), ("grid", 100, 3, "l1"), ("grid", 81, 4, "linf"), ("uniform", 200, 8, "l2"),
         ("gauss", 150, 3, "lp:3"), ("gauss", 2, 3, "l2"), ("gauss", 1, 3, "l2")]


@pytest.mark.parametrize("kind,n,dim,metric", CASES)
def test_vs_reference(harness, restatement, reference, kind, n, dim, metric):
    x = _cloud(kind, n, dim, n + 3 * dim)
    bars, _, _ = engine_diagram(harness, restatement, x, metric)
    got = canon(*bars)
    for mode in ("boundary", "coboundary"):
        res, _ = reference.persistence(x, maxdim=1, metric=metric, mode=mode)
        assert got == canon(*res), mode
    # with the ready-witness flags: no radius graph until witness searches
    # become common (the circles switch it on, the rest never need it)
    fbars, _, _ = engine_diagram(harness, restatement, x, metric, witness=True)
    assert canon(*fbars) == got


@pytest.mark.parametrize("kind,n,dim,metric", [("circle", 14, 2, "l2"), ("grid", 16, 2, "l1"),
                                                ("grid", 14, 3, "linf"), ("gauss", 12, 3, "l2")])
def test_witness_flags_brute_force(harness, restatement, kind, n, dim, metric):
    x = _cloud(kind, n, dim, 5 * n + dim)
    bars, mat, thr = engine_diagram(harness, restatement, x, metric, witness=True)
    assert canon(*bars) == canon(*brute_force(mat, thr))


def test_witness_flags_against_a_direct_loop(restatement):
    """The vectorised flag definition used above, against a plain loop over
    every w (the rule written out once more, case by case)."""
    x = _cloud("grid", 30, 2, 4)
    mat = restatement.make_dense(x, "l1")
    v, d = comb_edges(mat, restatement.enclosing_radius(mat, 30))
    got = witness_flags(mat, v, d)
    for (lo, hi), D, f in zip(v.tolist(), d.tolist(), got.tolist()):
        want = False
        for w in range(30):
            if w in (lo, hi) or mat[lo, w] > D or mat[hi, w] > D:
                continue
            if w < lo or (lo < w < hi and mat[hi, w] < D) or (w > hi and mat[lo, w] < D and mat[hi, w] < D):
                want = True
        assert f == want


def test_wrong_flags_change_the_diagram(harness, restatement):
    """A planted mistake in the flags (every flag cleared: no edge seems to
    have a witness) leaves classes alive that a triangle kills, and the
    brute force sees it."""
    x = _cloud("gauss", 14, 2, 8)
    mat = restatement.make_dense(x, "l2")
    thr = restatement.enclosing_radius(mat, 14)
    v, d = comb_edges(mat, thr)
    want = canon(*brute_force(mat, thr))
    assert canon(*engine_bars(harness, mat, with_flags(v, witness_flags(mat, v, d)), d, True)) == want
    assert canon(*engine_bars(harness, mat, with_flags(v, np.zeros(len(d), bool)), d, True)) != want


def test_ties_in_one_diameter(harness, restatement, reference):
    """A 4 x 4 x 4 integer grid under L-inf: each diameter class holds
    hundreds of equal edges, so the ready test's tie-breaks decide every
    triangle."""
    g = np.stack(np.meshgrid(*[np.arange(4)] * 3, indexing="ij"), -1).reshape(-1, 3).astype(np.float32)
    for metric in ("linf", "l1", "l2"):
        bars, _, _ = engine_diagram(harness, restatement, g, metric)
        res, _ = reference.persistence(g, maxdim=1, metric=metric)
        assert canon(*bars) == canon(*res), metric


def test_brute_force_catches_a_wrong_tie_break(harness, restatement):
    """The brute force is strong enough to see a mistake in the edge order:
    handing the engine the edges with ties in descending comb order (the
    coboundary engine's order) breaks the ready test on a tie-heavy grid."""
    x = _cloud("grid", 16, 2, 3)
    mat = restatement.make_dense(x, "l1")
    thr = restatement.enclosing_radius(mat, 16)
    v, d = comb_edges(mat, thr)
    want = canon(*brute_force(mat, thr))
    assert canon(*engine_bars(harness, mat, v, d)) == want
    n = 16
    hi, lo = np.tril_indices(n, -1)
    dd = mat[hi, lo]
    keep = dd <= thr
    hi, lo, dd = hi[keep], lo[keep], dd[keep]
    o = np.lexsort((-(hi.astype(np.int64) * (hi - 1) // 2 + lo), dd))
    bad_v = np.stack([lo[o], hi[o]], 1).astype(np.uint32)
    bad = canon(*engine_bars(harness, mat, bad_v, np.ascontiguousarray(dd[o], np.float32)))
    assert bad != want


# ---- the engine's set and map primitives, directly ----------------------------
@pytest.fixture(scope="module")
def prims():
    path = os.path.join(ROOT, "tests", "libhost_harness.so")
    if not os.path.exists(path):
        pytest.skip("tests/libhost_harness.so not built (make -C tests)")
    L = C.CDLL(path)
    L.hh_xor_into.argtypes = [_u32p, C.c_uint64, _u32p, C.c_uint64, _u32p]
    L.hh_xor_into.restype = C.c_uint64
    L.hh_rankmap.argtypes = [_u32p, C.c_uint64, _u32p]
    return L


@pytest.mark.parametrize("nx,nv", [(0, 1), (5, 1), (3000, 1), (4000, 3), (4000, 250), (100, 7), (50, 40),
                                   (0, 9), (17, 0), (1600, 100), (1601, 100)])
def test_xor_into_matches_set_symmetric_difference(prims, nx, nv):
    """All three branches of xor_into (one toggle; a few elements into a
    large set by binary search and run copies, v.size * 16 <= x.size; the
    merge), with members present and absent, against Python's set ^."""
    rng = np.random.default_rng(nx * 7 + nv)
    for trial in range(20):
        span = 3 * max(nx, 1) + 2 * nv
        x = np.sort(rng.choice(span, size=nx, replace=False)).astype(np.uint32)
        absent = np.setdiff1d(np.arange(span), x)
        take_in = min(nv // 2, x.size)
        v = np.unique(np.concatenate([rng.choice(x, size=take_in, replace=False) if take_in else x[:0],
                                      rng.choice(absent, size=nv - take_in, replace=False)])).astype(np.uint32)
        out = np.zeros(x.size + v.size + 1, np.uint32)
        k = prims.hh_xor_into(x.ctypes.data_as(_u32p), x.size, v.ctypes.data_as(_u32p), v.size,
                              out.ctypes.data_as(_u32p))
        assert out[:k].tolist() == sorted(set(x.tolist()) ^ set(v.tolist()))


def test_rankmap_matches_dict(prims):
    """2M random puts, gets and erases (erase by backward shift, growth by
    rehash) against a dict, keys > 0 in a narrow range so probe runs collide."""
    rng = np.random.default_rng(8)
    count = 2_000_000
    op = rng.choice([0, 1, 2], size=count, p=[0.45, 0.35, 0.2]).astype(np.uint32)
    key = rng.integers(1, 40000, size=count).astype(np.uint32)
    val = rng.integers(1, 2**32 - 1, size=count, dtype=np.uint64).astype(np.uint32)
    ops = np.ascontiguousarray(np.stack([op, key, val], 1))
    res = np.zeros(count, np.uint32)
    prims.hh_rankmap(ops.ctypes.data_as(_u32p), count, res.ctypes.data_as(_u32p))
    ref = {}
    want = np.zeros(count, np.uint32)
    for t in range(count):
        o, k = int(op[t]), int(key[t])
        if o == 0:
            ref[k] = int(val[t])
        elif o == 1:
            want[t] = ref.get(k, 0)
        else:
            ref.pop(k, None)
    assert np.array_equal(res, want)


@pytest.mark.parametrize("metric", ["l2", "l1", "linf"])
def test_sixteen_dimensional_clouds_vs_reference(harness, restatement, reference, metric):
    """16-D clouds, where the engine's resolves sum classes with many
    generators (the binary-search branch of xor_into, generator-free slots,
    G-set compaction): bar for bar against the reference, witness flags on
    and off."""
    x = np.random.default_rng({"l2": 71, "l1": 72, "linf": 73}[metric]).uniform(size=(900, 16)).astype(np.float32)
    want, _ = reference.persistence(x, maxdim=1, metric=metric)
    for witness in (False, True):
        got, _, _ = engine_diagram(harness, restatement, x, metric, witness)
        assert canon(*got) == canon(*want)
