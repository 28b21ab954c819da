"""Front-end operations of the reference (dnbaker/turbotops) on MI355X.

Host-array forms mirror the reference's C++ functions (same names, argument
meaning and errors); ``*_dev`` forms take torch device tensors and run on
the current HIP stream without leaving HBM. Every call goes through the C
ABI of libttgpu (include/ttgpu.h); nothing here computes on the CPU.
"""
import ctypes as C
import math

import numpy as np

from . import capi

_f32p = C.POINTER(C.c_float)


def _addr(a):
    return C.c_void_p(a.ctypes.data) if a.size else None


def _f32(x):
    return np.ascontiguousarray(x, dtype=np.float32)


def parse_metric(metric):
    """Metric::parse (include/turbotops/metric.hpp:240-253) -> (kind, p)."""
    k, p = C.c_int(), C.c_float()
    capi.lib().call("tt_metric_parse", str(metric).encode(), C.byref(k), C.byref(p))
    return k.value, p.value


def _metric(metric):
    if isinstance(metric, tuple):
        return metric
    return parse_metric(metric)


def dense_layout(n, budget):
    """make_dense's layout gate (rips.hpp:395-414): 'full', 'lower' or ValueError."""
    lay = C.c_int()
    capi.lib().call("tt_dense_layout", int(n), int(budget), C.byref(lay))
    return "full" if lay.value == capi.LAYOUT_FULL else "lower"


def make_dense(x, metric="l2", budget=2**64 - 1, with_radius=True):
    """rips::make_dense + DenseComplex::enclosing_radius (rips.hpp:370-416).

    Returns (full, mat, radius): mat is (n, n) for the full layout or the
    flat n(n-1)/2 lower triangle, written exactly as the reference lays it
    out; radius is the enclosing radius (None unless with_radius)."""
    x = _f32(x)
    if x.ndim != 2:
        raise ValueError("x must be a 2-d array (n, d)")
    n, d = x.shape
    kind, p = _metric(metric)
    layout = dense_layout(n, budget)
    full = layout == "full"
    cells = n * n if full else n * (n - 1) // 2
    mat = np.empty(cells, np.float32)
    r = C.c_float(math.inf)
    capi.lib().call("tt_make_dense", _addr(x), n, d, kind, p,
                    capi.LAYOUT_FULL if full else capi.LAYOUT_LOWER, _addr(mat),
                    C.byref(r) if with_radius else None)
    return full, (mat.reshape(n, n) if full else mat), (r.value if with_radius else None)


def _stream(stream):
    """A stream argument as a hipStream_t: None (torch's current stream), a
    torch.cuda.Stream, or a raw handle."""
    if stream is None:
        import torch
        return C.c_void_p(torch.cuda.current_stream().cuda_stream)
    if hasattr(stream, "cuda_stream"):
        return C.c_void_p(stream.cuda_stream)
    return C.c_void_p(int(stream))


def _dptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None and t.numel() else None


def _event_handle(ev, stream):
    """The hipEvent_t behind a torch.cuda.Event (torch creates it on the
    first record, so an event never recorded is recorded once here)."""
    if ev is None:
        return None
    if not ev.cuda_event:
        ev.record(stream)
    return C.c_void_p(ev.cuda_event)


def dist_matrix_dev(x, metric="l2", layout="full", out=None, rowmax=None, radius=None, stream=None,
                    events=None):
    """Device form: x is a (n, d) float32 torch tensor in HBM. Returns the
    matrix tensor (n*n or n(n-1)/2 floats); fills rowmax (n int32 holding
    float bits) and radius (1 float) when given. events: optional pair of
    torch.cuda.Event recorded right before and after the distance-tile
    kernel alone (stream must then be a torch.cuda.Stream)."""
    import torch
    if x.dtype != torch.float32 or not x.is_contiguous() or x.dim() != 2:
        raise ValueError("x must be a contiguous (n, d) float32 tensor")
    n, d = x.shape
    kind, p = _metric(metric)
    lay = capi.LAYOUT_FULL if layout == "full" else capi.LAYOUT_LOWER
    cells = n * n if lay == capi.LAYOUT_FULL else n * (n - 1) // 2
    if out is None:
        out = torch.empty(cells, dtype=torch.float32, device=x.device)
    elif out.numel() < cells:
        raise ValueError("out is too small")
    if events is None:
        capi.lib().call("tt_dist_matrix_dev", _dptr(x), n, d, kind, p, lay, _dptr(out), _dptr(rowmax),
                        _dptr(radius), _stream(stream))
    else:
        import torch
        ts = stream if isinstance(stream, torch.cuda.Stream) else torch.cuda.current_stream(x.device)
        ev0, ev1 = (_event_handle(e, ts) for e in events)
        capi.lib().call("tt_dist_matrix_dev_events", _dptr(x), n, d, kind, p, lay, _dptr(out), _dptr(rowmax),
                        _dptr(radius), C.c_void_p(ts.cuda_stream), ev0, ev1)
    return out


def _take(lib, ptr, count, ctype, dtype):
    """Copy a library-malloc'd host buffer into numpy and release it."""
    if not ptr:
        return np.zeros(0, dtype)
    arr = np.ctypeslib.as_array(ptr, shape=(max(count, 1),))[:count].astype(dtype, copy=True)
    lib.dll.tt_free(C.cast(ptr, C.c_void_p))
    return arr


def dense_edges(x, metric="l2", threshold=math.inf, apparent=True):
    """Reducer::dim0's kept edges, sort and apparent test on the dense full
    complex (rips.hpp:685-849). Returns (idx u64, d f32, apparent u8 or None,
    threshold), edges ordered by (diameter asc, index desc)."""
    x = _f32(x)
    n, d = x.shape
    kind, p = _metric(metric)
    lib = capi.lib()
    m = C.c_uint64()
    pi, pd, pa = C.POINTER(C.c_uint64)(), C.POINTER(C.c_float)(), C.POINTER(C.c_uint8)()
    thr = C.c_float()
    lib.call("tt_dense_edges", _addr(x), n, d, kind, p, C.c_float(threshold), C.byref(m), C.byref(pi),
             C.byref(pd), C.byref(pa) if apparent else None, C.byref(thr))
    k = m.value
    idx = _take(lib, pi, k, C.c_uint64, np.uint64)
    dd = _take(lib, pd, k, C.c_float, np.float32)
    ap = _take(lib, pa, k, C.c_uint8, np.uint8) if apparent else None
    return idx, dd, ap, thr.value


def dense_edges_dev(mat, threshold, apparent=True, count_only=False, stream=None):
    """Device form of dense_edges over a full (n, n) float32 matrix already in
    HBM (a torch tensor): (m, idx, d, apparent) as torch tensors on the same
    device, or just m with count_only (the whole pipeline still runs)."""
    import torch
    n = int(mat.shape[0])
    if mat.dtype != torch.float32 or not mat.is_contiguous() or mat.numel() != n * n:
        raise ValueError("mat must be a contiguous float32 (n, n) tensor")
    lib = capi.lib()
    m = C.c_uint64()
    st = _stream(stream)
    lib.call("tt_dense_edges_dev", C.c_void_p(mat.data_ptr()), n, C.c_float(threshold), int(bool(apparent)), 0,
             C.byref(m), None, None, None, st)
    k = m.value
    if count_only:
        return k
    idx = torch.empty(k, dtype=torch.int64, device=mat.device)
    dd = torch.empty(k, dtype=torch.float32, device=mat.device)
    ap = torch.empty(k, dtype=torch.uint8, device=mat.device) if apparent else None
    lib.call("tt_dense_edges_dev", C.c_void_p(mat.data_ptr()), n, C.c_float(threshold), int(bool(apparent)), k,
             C.byref(m), C.c_void_p(idx.data_ptr()), C.c_void_p(dd.data_ptr()),
             C.c_void_p(ap.data_ptr()) if apparent else None, st)
    return k, idx, dd, ap


def sparse_edges(graph, threshold=math.inf, apparent=True):
    """Reducer::dim0's kept edges, sort and apparent test on the flag complex
    of a CSR graph (rips.hpp:685-849, SparseComplex::cofacets :428-471).
    graph: a CsrGraph or (n, row_ptr, col, w). Returns (idx u64, d f32,
    apparent u8 or None), edges ordered by (diameter asc, index desc)."""
    if isinstance(graph, CsrGraph):
        n, rp, col, w = graph.n, graph.row_ptr, graph.col_idx, graph.weights
    else:
        n, rp, col, w = graph
    rp = np.ascontiguousarray(rp, np.uint64)
    col = np.ascontiguousarray(col, np.uint32)
    w = np.ascontiguousarray(w, np.float32)
    if rp.shape != (int(n) + 1,):
        raise ValueError("row_ptr must have n + 1 entries")
    if col.shape[0] < int(rp[-1]) or w.shape[0] < int(rp[-1]):
        raise ValueError("col and w must hold row_ptr[n] entries")
    lib = capi.lib()
    m = C.c_uint64()
    pi, pd, pa = C.POINTER(C.c_uint64)(), C.POINTER(C.c_float)(), C.POINTER(C.c_uint8)()
    lib.call("tt_sparse_edges", int(n), _addr(rp), _addr(col), _addr(w), C.c_float(threshold), C.byref(m),
             C.byref(pi), C.byref(pd), C.byref(pa) if apparent else None)
    k = m.value
    idx = _take(lib, pi, k, C.c_uint64, np.uint64)
    dd = _take(lib, pd, k, C.c_float, np.float32)
    ap = _take(lib, pa, k, C.c_uint8, np.uint8) if apparent else None
    return idx, dd, ap


def build_edges(x, metric="l2", threshold=math.inf, layout="full", with_matrix=False, witness=False):
    """The boundary (RedZeD) engine's front end: make_dense + enclosing
    radius (turbotops.hpp:243-246) and build_edges (redzed.hpp:544-613).
    Returns (verts (m, 2) int32 with verts[:, 0] < verts[:, 1], diam f32,
    threshold[, mat]) in (diameter asc, combinatorial index asc) order.
    witness=True (tt_build_edges_witness) adds, after diam, each edge's
    ready-witness flag (bool): a triangle exists whose youngest face in this
    order is the edge."""
    x = _f32(x)
    if x.ndim != 2:
        raise ValueError("x must be a 2-d array (n, d)")
    n, dim = x.shape
    kind, p = _metric(metric)
    lib = capi.lib()
    full = layout == "full"
    mat = None
    if with_matrix:
        mat = np.empty(n * n if full else n * (n - 1) // 2, np.float32)
    m = C.c_uint64()
    pv, pd = C.POINTER(C.c_int32)(), C.POINTER(C.c_float)()
    thr = C.c_float()
    lib.call("tt_build_edges_witness" if witness else "tt_build_edges", _addr(x), n, dim, kind, p,
             C.c_float(threshold), capi.LAYOUT_FULL if full else capi.LAYOUT_LOWER,
             _addr(mat) if mat is not None else None, C.byref(m), C.byref(pv), C.byref(pd), C.byref(thr))
    k = m.value
    verts = _take(lib, pv, 2 * k, C.c_int32, np.int32).reshape(k, 2)
    diam = _take(lib, pd, k, C.c_float, np.float32)
    out = [verts, diam]
    if witness:
        u = verts.view(np.uint32)
        out.append((u[:, 1] >> 31).astype(bool))
        verts[:, 1] = (u[:, 1] & 0x7FFFFFFF).astype(np.int32)
    out.append(thr.value)
    if with_matrix:
        out.append(mat.reshape(n, n) if full else mat)
    return tuple(out)


def build_edges_csr(graph, threshold=math.inf):
    """build_edges over a CSR graph (redzed.hpp:544-613 with SparseGeom):
    (verts (m, 2) int32, diam f32) in (diameter asc, combinatorial index asc)
    order."""
    if isinstance(graph, CsrGraph):
        n, rp, col, w = graph.n, graph.row_ptr, graph.col_idx, graph.weights
    else:
        n, rp, col, w = graph
    rp = np.ascontiguousarray(rp, np.uint64)
    col = np.ascontiguousarray(col, np.uint32)
    w = np.ascontiguousarray(w, np.float32)
    if rp.shape != (int(n) + 1,):
        raise ValueError("row_ptr must have n + 1 entries")
    if col.shape[0] < int(rp[-1]) or w.shape[0] < int(rp[-1]):
        raise ValueError("col and w must hold row_ptr[n] entries")
    lib = capi.lib()
    m = C.c_uint64()
    pv, pd = C.POINTER(C.c_int32)(), C.POINTER(C.c_float)()
    lib.call("tt_build_edges_csr", int(n), _addr(rp), _addr(col), _addr(w), C.c_float(threshold), C.byref(m),
             C.byref(pv), C.byref(pd))
    k = m.value
    return _take(lib, pv, 2 * k, C.c_int32, np.int32).reshape(k, 2), _take(lib, pd, k, C.c_float, np.float32)


class CsrGraph:
    """CsrGraph (include/turbotops/core.hpp:92-149): n vertices, m undirected
    edges stored in both rows, rows sorted by neighbour id, no self loops."""

    def __init__(self, n, row_ptr, col_idx, weights):
        self.n = int(n)
        self.row_ptr = np.asarray(row_ptr, np.uint64)
        self.col_idx = np.asarray(col_idx, np.uint32)
        self.weights = np.asarray(weights, np.float32)
        self.m = int(self.row_ptr[-1]) // 2 if self.row_ptr.size else 0

    def valid(self):
        """CsrGraph::valid (core.hpp:137-148)."""
        n = self.n
        if self.row_ptr.size != n + 1 or self.row_ptr[0] != 0 or int(self.row_ptr[n]) != 2 * self.m:
            return False
        for v in range(n):
            a, b = int(self.row_ptr[v]), int(self.row_ptr[v + 1])
            cols = self.col_idx[a:b]
            if np.any(cols >= n) or np.any(cols == v) or np.any(np.diff(cols.astype(np.int64)) <= 0):
                return False
            for t, u in enumerate(cols):
                ua, ub = int(self.row_ptr[u]), int(self.row_ptr[u + 1])
                k = np.searchsorted(self.col_idx[ua:ub], v)
                if k >= ub - ua or self.col_idx[ua + k] != v or self.weights[ua + k] != self.weights[a + t]:
                    return False
        return True


def _csr_call(name, n, *args):
    lib = capi.lib()
    rp = np.zeros(n + 1, np.uint64)
    nnz = C.c_uint64()
    pc, pw = C.POINTER(C.c_uint32)(), C.POINTER(C.c_float)()
    lib.call(name, *args, _addr(rp) if rp.size else None, C.byref(nnz), C.byref(pc), C.byref(pw))
    k = nnz.value
    return CsrGraph(n, rp, _take(lib, pc, k, C.c_uint32, np.uint32), _take(lib, pw, k, C.c_float, np.float32))


def radius_graph(x, metric, tau):
    """radius_graph(PointCloud, tau) (graph.hpp:170-201)."""
    x = _f32(x)
    n, d = x.shape
    kind, p = _metric(metric)
    return _csr_call("tt_radius_graph", n, _addr(x), n, d, kind, p, C.c_float(tau))


def knn_lists_exact(x, metric, k):
    """knn_lists_exact(PointCloud, k) (graph.hpp:211-225): (n, min(k, n-1)) ids."""
    x = _f32(x)
    n, d = x.shape
    kind, p = _metric(metric)
    kk = min(k, n - 1) if n else 0
    ids = np.zeros(max(n * kk, 1), np.uint32)
    kout = C.c_uint32()
    capi.lib().call("tt_knn_exact", _addr(x), n, d, kind, p, int(k), C.byref(kout), _addr(ids))
    return ids[: n * kout.value].reshape(n, kout.value)


def knn_graph(x, metric, ids):
    """knn_graph(PointCloud, KnnLists) (graph.hpp:286-304)."""
    x = _f32(x)
    n, d = x.shape
    kind, p = _metric(metric)
    ids = np.ascontiguousarray(ids, np.uint32)
    k = ids.shape[1] if ids.ndim == 2 else 0
    return _csr_call("tt_knn_graph", n, _addr(x), n, d, kind, p, k, _addr(ids) if ids.size else None)


def knn_graph_exact(x, metric, k):
    """knn_graph(cloud, knn_lists_exact(cloud, k)) with the lists kept in HBM."""
    x = _f32(x)
    n, d = x.shape
    kind, p = _metric(metric)
    return _csr_call("tt_knn_graph_exact", n, _addr(x), n, d, kind, p, int(k))


def mst_deaths(x, metric="l2"):
    """The n - 1 sorted H0 deaths (MST edge weights) by Boruvka."""
    x = _f32(x)
    n, d = x.shape
    kind, p = _metric(metric)
    out = np.zeros(max(n - 1, 1), np.float32)
    capi.lib().call("tt_mst_deaths", _addr(x), n, d, kind, p, _addr(out))
    return out[: max(n - 1, 0)]


def mst_h0(x, metric="l2", threshold=math.inf):
    """mst::h0(PointCloud, threshold) (mst.hpp:25-99): H0 bars (births, deaths)
    in the reference's order: finite deaths ascending, then the essential
    classes (one, plus one per MST edge past the threshold)."""
    deaths = mst_deaths(x, metric)
    n = _f32(x).shape[0]
    if n == 0:
        return np.zeros(0, np.float32), np.zeros(0, np.float32)
    thr = np.float32(threshold)
    essential = 1 + int(np.count_nonzero(deaths > thr))
    fin = deaths[(deaths <= thr) & (deaths > 0)]
    d = np.concatenate([fin, np.full(essential, np.inf, np.float32)]).astype(np.float32)
    return np.zeros_like(d), d


def _words(values, words):
    vals = [int(v) for v in values]
    if words == 1:
        return np.array(vals, dtype=np.uint64)
    return np.array([[v & (2**64 - 1), v >> 64] for v in vals], dtype=np.uint64).reshape(-1, 2)


def _values(arr, words):
    a = np.asarray(arr, dtype=np.uint64)
    if words == 1:
        return [int(v) for v in a.ravel()]
    return [int(lo) | (int(hi) << 64) for lo, hi in a.reshape(-1, 2)]


class DeviceComplex:
    """The coboundary reducer's complex in HBM (include/ttgpu.h tt_complex_*):
    the per-simplex work above dimension 0 — emergent candidates of columns
    (rips.hpp:903-928) and the next dimension's columns (Reducer::assemble,
    :981-1023). dense: an (n, n) float32 matrix (full layout) or the flat
    lower triangle; sparse: a CsrGraph. Indices are Python ints; words is
    the index width in 64-bit words (1 or 2)."""

    def __init__(self, dense=None, graph=None, threshold=math.inf, maxdim=2, words=1):
        lib = capi.lib()
        self.words = int(words)
        h = C.c_void_p()
        if dense is not None:
            m = _f32(dense)
            full = m.ndim == 2
            n = m.shape[0] if full else int((1 + math.isqrt(1 + 8 * m.size)) // 2)
            self._keep = m
            lib.call("tt_complex_dense", _addr(m), n, capi.LAYOUT_FULL if full else capi.LAYOUT_LOWER, 0,
                     float(threshold), int(maxdim), self.words, C.byref(h))
        else:
            g = graph
            self._keep = g
            lib.call("tt_complex_sparse", g.n, _addr(g.row_ptr), _addr(g.col_idx), _addr(g.weights),
                     float(threshold), int(maxdim), self.words, C.byref(h))
        self._h = h

    def close(self):
        if self._h:
            capi.lib().dll.tt_complex_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def emergent(self, nv, idx, d):
        """(cofacets, states) per column: state 1 a candidate emergent pair,
        2 its cofacet has a zero-apparent facet, 0 none."""
        w = self.words
        iw = _words(idx, w)
        d = _f32(d)
        cof = np.zeros((max(len(d), 1), w), np.uint64)
        st = np.zeros(max(len(d), 1), np.uint8)
        capi.lib().call("tt_complex_emergent", self._h, int(nv), _addr(iw), _addr(d), len(d), _addr(cof),
                        _addr(st))
        return _values(cof[: len(d)], w), st[: len(d)]

    def seed(self, nv, idx, d):
        iw = _words(idx, self.words)
        d = _f32(d)
        capi.lib().call("tt_complex_seed", self._h, int(nv), _addr(iw), _addr(d), len(d))

    def assemble(self, pivots=(), keep_next=False):
        """(columns [(idx, d)], (cofacets, states), skipped)."""
        lib = capi.lib()
        w = self.words
        pv = _words(pivots, w)
        nc, sk = C.c_uint64(), C.c_uint64()
        pi, pc = C.POINTER(C.c_uint64)(), C.POINTER(C.c_uint64)()
        pd = _f32p()
        ps = C.POINTER(C.c_uint8)()
        lib.call("tt_complex_assemble", self._h, _addr(pv), len(pv), int(bool(keep_next)), C.byref(nc),
                 C.byref(pi), C.byref(pd), C.byref(pc), C.byref(ps), C.byref(sk))
        k = nc.value
        idx = _values(_take(lib, pi, k * w, C.c_uint64, np.uint64), w)
        d = _take(lib, pd, k, C.c_float, np.float32)
        cof = _values(_take(lib, pc, k * w, C.c_uint64, np.uint64), w)
        st = _take(lib, ps, k, C.c_uint8, np.uint8)
        return list(zip(idx, d.tolist())), (cof, st), sk.value
