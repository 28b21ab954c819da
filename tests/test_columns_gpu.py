"""The coboundary reducer's per-simplex work above dimension 0 on the GPU
(include/ttgpu.h tt_complex_*, csrc/columns.hip) against the oracle's
restatement (tto_emergent_cx, tto_assemble_cx, pinned against the
reference's Reducer in test_oracle_cx.py), bit for bit: every column index
and diameter in order, every emergent state and cofacet, the skipped count,
with 64- and 128-bit indices, on dense and sparse complexes including
tie-heavy grids; and end to end, diagrams with 128-bit indices forced equal
the 64-bit ones, and a natural 128-bit case (C(150000, 4) > 2^64) matches
the compiled reference bar for bar."""
import math

import numpy as np
import pytest

from test_oracle_cx import dense_cases, edges_of, sparse_cases

pytestmark = pytest.mark.gpu


def _device(cx, thr, words, maxdim=3):
    from turbotops_amd import CsrGraph, DeviceComplex
    if not cx.sparse:
        return DeviceComplex(dense=cx.mat, threshold=thr, maxdim=maxdim, words=words)
    g = CsrGraph(cx.n, cx.rp, cx.col, cx.w)
    return DeviceComplex(graph=g, threshold=thr, maxdim=maxdim, words=words)


def _same_cols(a, b):
    return [(i, np.float32(d)) for i, d in a] == [(i, np.float32(d)) for i, d in b]


@pytest.mark.parametrize("words", [1, 2])
def test_emergent_matches_oracle(gpu, restatement, words):
    for name, cx, thr in dense_cases(restatement) + sparse_cases(restatement):
        idx, d = edges_of(restatement, cx, thr)
        idx = [int(v) for v in idx]
        print(f"emergent case {name} words {words} edges {len(idx)}", flush=True)
        dev = _device(cx, thr, words)
        cof, st = dev.emergent(2, idx, d)
        wcof, wst = restatement.emergent_cx(cx, thr, 2, idx, d, words)
        assert np.array_equal(st, wst), name
        assert [c for c, s in zip(cof, st) if s] == [c for c, s in zip(wcof, wst) if s], name
        # triangles (from the oracle's assemble) as columns
        _, _, tri, _ = restatement.assemble_cx(cx, thr, 2, idx, d, [], True, words)
        ti = [v for v, _ in tri][:3000]
        td = [x for _, x in tri][:3000]
        if ti:
            cof, st = dev.emergent(3, ti, td)
            wcof, wst = restatement.emergent_cx(cx, thr, 3, ti, td, words)
            assert np.array_equal(st, wst), name
            assert [c for c, s in zip(cof, st) if s] == [c for c, s in zip(wcof, wst) if s], name
        dev.close()


@pytest.mark.parametrize("words", [1, 2])
def test_assemble_matches_oracle(gpu, restatement, words):
    for name, cx, thr in dense_cases(restatement) + sparse_cases(restatement):
        idx, d = edges_of(restatement, cx, thr)
        idx = [int(v) for v in idx]
        _, _, tri, _ = restatement.assemble_cx(cx, thr, 2, idx, d, [], True, words)
        piv = [v for v, _ in tri[::3]]
        print(f"assemble case {name} words {words} edges {len(idx)} triangles {len(tri)}", flush=True)
        dev = _device(cx, thr, words)
        dev.seed(2, idx, d)
        cols, (cof, st), sk = dev.assemble(piv, keep_next=True)
        wcols, (wcof, wst), wnext, wsk = restatement.assemble_cx(cx, thr, 2, idx, d, piv, True, words)
        assert _same_cols(cols, wcols), name
        assert np.array_equal(st, wst) and sk == wsk, name
        assert [c for c, s in zip(cof, st) if s] == [c for c, s in zip(wcof, wst) if s], name
        # the cofacets stayed on the device: dimension 3 from them
        cols3, (cof3, st3), sk3 = dev.assemble([], keep_next=False)
        wi = [v for v, _ in wnext]
        wd = [x for _, x in wnext]
        w3, (wc3, ws3), _, wsk3 = restatement.assemble_cx(cx, thr, 3, wi, wd, [], False, words)
        assert _same_cols(cols3, w3), name
        assert np.array_equal(st3, ws3) and sk3 == wsk3, name
        assert [c for c, s in zip(cof3, st3) if s] == [c for c, s in zip(wc3, ws3) if s], name
        dev.close()


def test_lower_layout_complex_equals_full(gpu, restatement):
    from turbotops_amd import DeviceComplex
    name, cx, thr = dense_cases(restatement)[1]
    n = cx.n
    low = np.array([cx.mat[i, j] for i in range(n) for j in range(i)], np.float32)
    idx, d = edges_of(restatement, cx, thr)
    idx = [int(v) for v in idx]
    a = DeviceComplex(dense=cx.mat, threshold=thr, maxdim=2)
    b = DeviceComplex(dense=low, threshold=thr, maxdim=2)
    a.seed(2, idx, d)
    b.seed(2, idx, d)
    ca, sa, ka = a.assemble()
    cb, sb, kb = b.assemble()
    assert _same_cols(ca, cb) and np.array_equal(sa[1], sb[1]) and ka == kb
    assert np.array_equal(a.emergent(2, idx, d)[1], b.emergent(2, idx, d)[1])


def test_bad_arguments_raise(gpu):
    from turbotops_amd import DeviceComplex
    with pytest.raises(ValueError, match="index_words"):
        DeviceComplex(dense=np.zeros((4, 4), np.float32), threshold=1.0, words=3)
    with pytest.raises(OverflowError, match="simplex index overflow"):
        DeviceComplex(graph=_line_graph(300_000), threshold=1.0, maxdim=2, words=1)   # C(3e5, 4) > 2^64
    dev = DeviceComplex(dense=np.zeros((4, 4), np.float32), threshold=1.0, maxdim=1)
    with pytest.raises(ValueError, match="no simplices seeded"):
        dev.assemble()


def _line_graph(n):
    from turbotops_amd import CsrGraph
    rp = np.zeros(n + 1, np.uint64)
    return CsrGraph(n, rp, np.zeros(0, np.uint32), np.zeros(0, np.float32))


def _canon(dgms):
    out = []
    for k, a in enumerate(dgms):
        for b, e in np.asarray(a).tolist():
            out.append((k, b, e))
    return sorted(out)


@pytest.mark.parametrize("case", ["dense-torus", "dense-grid-l1", "knn"])
def test_forced_wide_diagrams_equal_narrow(gpu, case):
    """persistence with 128-bit indices forced (the reference's force_wide,
    rips.hpp:1038-1061; its self-test src/turbotops.cpp:446-450) gives the
    64-bit diagrams bar for bar, through the GPU column stages."""
    from turbotops_amd import persistence
    rng = np.random.default_rng(44)
    if case == "dense-torus":
        a, b = rng.uniform(0, 2 * np.pi, size=(2, 160))
        x = np.stack([(2 + np.cos(a)) * np.cos(b), (2 + np.cos(a)) * np.sin(b), np.sin(a)], 1).astype(np.float32)
        kw = dict(maxdim=2)
    elif case == "dense-grid-l1":
        x = rng.integers(0, 4, size=(60, 3)).astype(np.float32)
        kw = dict(maxdim=3, metric="l1")
    else:
        x = rng.uniform(size=(2000, 5)).astype(np.float32)
        kw = dict(maxdim=2, graph="knn", k=10)
    print(f"forced-wide case {case}", flush=True)
    narrow, i1 = persistence(x, _mode="coboundary", **kw)
    wide, i2 = persistence(x, _mode="coboundary", _force_wide=True, **kw)
    assert i1["index_bits"] == 64 and i2["index_bits"] == 128
    assert _canon(narrow) == _canon(wide)
    assert sum(len(a) for a in narrow[1:]) > 0


def test_natural_128_bit_case_vs_reference(gpu, reference):
    """150000 points uniform in [0,1]^8, radius graph at tau = 0.25, maxdim 2:
    C(150000, 4) > 2^64, so both engines run on 128-bit indices; bar for bar
    against the compiled reference (about 129K H1 and 77 H2 bars)."""
    from turbotops_amd import persistence
    x = np.random.default_rng(5).uniform(size=(150000, 8)).astype(np.float32)
    print("natural u128 case n=150000 d=8 tau=0.25", flush=True)
    dg, info = persistence(x, maxdim=2, threshold=0.25)
    assert info["index_bits"] == 128
    (dims, b, e), _ = reference.persistence(x, maxdim=2, threshold=0.25)
    want = sorted(zip(dims.tolist(), b.tolist(), e.tolist()))
    assert _canon(dg) == want
    assert len(dg[2]) > 0
