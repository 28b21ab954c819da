"""Pins of the oracle's reducer tests on dense and sparse complexes
(oracle/ttoracle.c: tto_apparent_cx, tto_emergent_cx, tto_assemble_cx, and
through them tto_apparent_cofacet / tto_apparent_facet):

* against the reference's own Reducer members, compiled from its sources
  (oracle/_ref: zero_apparent_cofacet / zero_apparent_facet,
  rips.hpp:638-665; the emergent shortcut's first equal-diameter cofacet,
  :903-928; Reducer::assemble, :981-1023), with 64- and 128-bit indices;
* against a from-scratch Python statement of the definitions (largest w
  for the zero-pivot cofacet, first facet in ascending index order for the
  zero-pivot facet) on small complexes;
* and planted mistakes in that statement (the facet order reversed, the
  cofacet search ascending) are shown to disagree, so the checks can fail.
"""
from math import comb

import numpy as np
import pytest

from oracle import Complex, Restatement

pytestmark = []


def grid(side, dim):
    g = np.stack(np.meshgrid(*[np.arange(side)] * dim, indexing="ij"), -1).reshape(-1, dim)
    return g.astype(np.float32)


def dense_cases(R):
    rng = np.random.default_rng(21)
    out = []
    for name, x, metric in [("uniform2", rng.random((40, 2)), "l2"),
                            ("gauss3", rng.normal(size=(36, 3)), "l1"),
                            ("grid-l1", grid(4, 2), "l1"),
                            ("grid-linf", grid(3, 3), "linf"),
                            ("circles", np.r_[np.c_[np.cos(np.arange(20) * 0.314), np.sin(np.arange(20) * 0.314)],
                                              np.c_[2.5 + 0.5 * np.cos(np.arange(12) * 0.52),
                                                    0.5 * np.sin(np.arange(12) * 0.52)]], "l2")]:
        x = np.asarray(x, np.float32)
        mat = R.make_dense(x, metric, True)
        out.append((name, Complex(mat=mat), R.enclosing_radius(mat, x.shape[0], True)))
    return out


def sparse_cases(R):
    rng = np.random.default_rng(22)
    out = []
    x = rng.random((120, 3)).astype(np.float32)
    rp, col, w = R.knn_graph(x, "l2", R.knn(x, "l2", 6))
    out.append(("knn6", Complex(n=120, row_ptr=rp, col=col, w=w), np.inf))
    g = grid(4, 3)
    rp, col, w = R.radius_graph(g, "l1", 2.0)
    out.append(("grid-radius", Complex(n=g.shape[0], row_ptr=rp, col=col, w=w), np.inf))
    x = rng.random((90, 2)).astype(np.float32)
    rp, col, w = R.radius_graph(x, "linf", 0.3)
    out.append(("linf-radius", Complex(n=90, row_ptr=rp, col=col, w=w), 0.2))
    return out


def edges_of(R, cx, thr):
    """kept edges (idx, d) of a complex at threshold thr"""
    if not cx.sparse:
        return R.dense_edges(cx.mat, thr)
    idx, d = [], []
    for v in range(cx.n):
        for t in range(int(cx.rp[v]), int(cx.rp[v + 1])):
            u = int(cx.col[t])
            if u < v and cx.w[t] <= thr:
                idx.append(v * (v - 1) // 2 + u)
                d.append(cx.w[t])
    o = np.lexsort((-np.array(idx, dtype=np.float64), np.array(d, np.float32)))
    return np.array(idx, np.uint64)[o], np.array(d, np.float32)[o]


# ---- a from-scratch statement of the definitions --------------------------
def dist(cx, a, b):
    if a == b:
        return 0.0
    if not cx.sparse:
        return float(cx.mat[a, b])
    row = cx.col[int(cx.rp[a]):int(cx.rp[a + 1])]
    k = np.searchsorted(row, b)
    if k < len(row) and row[k] == b:
        return float(cx.w[int(cx.rp[a]) + k])
    return None


def index(vs):
    vs = sorted(vs, reverse=True)
    return sum(comb(v, len(vs) - i) for i, v in enumerate(vs))


def decode(idx, k, n):
    out, top = [], n
    for pos in range(k, 0, -1):
        v = top - 1
        while comb(v, pos) > idx:
            v -= 1
        out.append(v)
        idx -= comb(v, pos)
        top = v
    return out


def zpc(cx, vs, d, descending=True):
    ws = range(cx.n - 1, -1, -1) if descending else range(cx.n)
    for w in ws:
        if w in vs:
            continue
        ds = [dist(cx, v, w) for v in vs]
        if all(x is not None and x <= d for x in ds):
            return sorted(vs + [w], reverse=True)
    return None


def zpf(cx, vs, d, ascending=True):
    vs = sorted(vs, reverse=True)
    js = range(len(vs)) if ascending else range(len(vs) - 1, -1, -1)
    for j in js:
        f = vs[:j] + vs[j + 1:]
        if any(dist(cx, f[a], f[b]) == d for a in range(len(f)) for b in range(a + 1, len(f))):
            return f
    return None


def apparent_brute(cx, vs, d, planted=None):
    c = zpc(cx, vs, d, descending=planted != "cofacet-ascending")
    cof = c is not None and zpf(cx, c, d, ascending=planted != "facet-descending") == sorted(vs, reverse=True)
    f = zpf(cx, vs, d, ascending=planted != "facet-descending")
    fac = f is not None and zpc(cx, f, d, descending=planted != "cofacet-ascending") == sorted(vs, reverse=True)
    return int(cof) | (int(fac) << 1)


def simplices(R, cx, thr, nv, cap=4000):
    idx, d = edges_of(R, cx, thr)
    idx = [int(v) for v in idx]
    d = [float(v) for v in d]
    if nv == 2:
        return idx[:cap], d[:cap]
    from oracle import Reference
    Ref = Reference()
    cur = (idx, d)
    for k in range(2, nv):
        _, nxt, _ = Ref.assemble_cx(cx, thr, k, cur[0], cur[1], [], True)
        cur = ([v for v, _ in nxt], [x for _, x in nxt])
    return cur[0][:cap], cur[1][:cap]


@pytest.mark.parametrize("words", [1, 2])
@pytest.mark.parametrize("nv", [2, 3, 4])
def test_apparent_cx_matches_reference(restatement, reference, nv, words):
    for name, cx, thr in dense_cases(restatement) + sparse_cases(restatement):
        idx, d = simplices(restatement, cx, thr, nv)
        if not idx:
            continue
        got = restatement.apparent_cx(cx, nv, idx, d, words)
        want = reference.apparent_cx(cx, nv, idx, d, words)
        assert np.array_equal(got, want), (name, nv, words)


def test_apparent_facet_pinned_on_triangles(restatement, reference):
    """tto_apparent_facet (full matrix) against the reference's
    zero_apparent_facet and the brute force, on every triangle of tie-heavy
    grids and a random cloud; both outcomes occur."""
    seen = set()
    for name, cx, thr in dense_cases(restatement):
        idx, d = simplices(restatement, cx, thr, 3, cap=1500)
        want = reference.apparent_cx(cx, 3, idx, d) >> 1
        for t, (s, ds) in enumerate(zip(idx, d)):
            got = restatement.apparent_facet(cx.mat, 3, s, ds)
            assert got == bool(want[t]), (name, s)
            seen.add(got)
        ex = restatement.apparent_cx(cx, 3, idx, d) & 1
        for t, (s, ds) in enumerate(zip(idx, d)):
            assert restatement.apparent_cofacet(cx.mat, 3, s, ds) == bool(ex[t])
    assert seen == {True, False}


def test_apparent_cx_brute_force_and_planted(restatement):
    planted_caught = {"facet-descending": False, "cofacet-ascending": False}
    for name, cx, thr in dense_cases(restatement)[:4] + sparse_cases(restatement)[1:2]:
        for nv in (2, 3):
            idx, d = simplices(restatement, cx, thr, nv, cap=300)
            got = restatement.apparent_cx(cx, nv, idx, d)
            for t, (s, ds) in enumerate(zip(idx, d)):
                vs = decode(s, nv, cx.n)
                assert got[t] == apparent_brute(cx, vs, ds), (name, nv, vs)
                for p in planted_caught:
                    if apparent_brute(cx, vs, ds, planted=p) != got[t]:
                        planted_caught[p] = True
    assert all(planted_caught.values()), planted_caught


@pytest.mark.parametrize("words", [1, 2])
@pytest.mark.parametrize("nv", [2, 3])
def test_emergent_cx_matches_reference(restatement, reference, nv, words):
    states = set()
    for name, cx, thr in dense_cases(restatement) + sparse_cases(restatement):
        idx, d = simplices(restatement, cx, thr, nv, cap=2500)
        if not idx:
            continue
        c1, s1 = restatement.emergent_cx(cx, thr, nv, idx, d, words)
        c2, s2 = reference.emergent_cx(cx, thr, nv, idx, d, words)
        assert np.array_equal(s1, s2), (name, nv)
        assert [a for a, s in zip(c1, s1) if s] == [b for b, s in zip(c2, s2) if s], (name, nv)
        states |= set(s1.tolist())
    assert {0, 2} <= states


def test_emergent_cx_brute_force(restatement):
    for name, cx, thr in dense_cases(restatement)[:3]:
        idx, d = simplices(restatement, cx, thr, 2, cap=400)
        cof, st = restatement.emergent_cx(cx, thr, 2, idx, d)
        for t, (s, ds) in enumerate(zip(idx, d)):
            vs = decode(s, 2, cx.n)
            c = zpc(cx, vs, ds)
            if c is None:
                assert st[t] == 0
                continue
            assert cof[t] == index(c)
            blocked = apparent_brute(cx, c, ds) >> 1
            assert st[t] == (2 if blocked else 1), (name, vs)


@pytest.mark.parametrize("words", [1, 2])
def test_assemble_cx_matches_reference(restatement, reference, words):
    for name, cx, thr in dense_cases(restatement) + sparse_cases(restatement):
        idx, d = edges_of(restatement, cx, thr)
        idx = [int(v) for v in idx]
        d = [float(v) for v in d]
        # pivots: a deterministic third of the triangles, so clearing is exercised
        _, tri, _ = reference.assemble_cx(cx, thr, 2, idx, d, [], True, words)
        piv = [v for v, _ in tri[::3]]
        for keep in (0, 1):
            cols, em, nxt, sk = restatement.assemble_cx(cx, thr, 2, idx, d, piv, keep, words)
            rc, rn, rsk = reference.assemble_cx(cx, thr, 2, idx, d, piv, keep, words)
            assert [(a, np.float32(b)) for a, b in cols] == [(a, np.float32(b)) for a, b in rc], name
            assert sk == rsk
            if keep:
                assert nxt == rn
            ci = [a for a, _ in cols]
            cd = [b for _, b in cols]
            c2, s2 = reference.emergent_cx(cx, thr, 3, ci, cd, words)
            assert np.array_equal(em[1], s2)
            assert [a for a, s in zip(em[0], em[1]) if s] == [b for b, s in zip(c2, s2) if s]
        # dimension 3 columns from the triangles
        tri_i = [v for v, _ in tri][:1500]
        tri_d = [x for _, x in tri][:1500]
        cols, em, _, sk = restatement.assemble_cx(cx, thr, 3, tri_i, tri_d, [], 0, words)
        rc, _, rsk = reference.assemble_cx(cx, thr, 3, tri_i, tri_d, [], 0, words)
        assert [(a, np.float32(b)) for a, b in cols] == [(a, np.float32(b)) for a, b in rc], name
        assert sk == rsk
