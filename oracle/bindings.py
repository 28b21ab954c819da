"""ctypes bindings of the oracle libraries (test infrastructure only)."""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
METRIC_KINDS = {"l1": 0, "l2": 1, "linf": 2, "lp": 3}

_f32p = C.POINTER(C.c_float)
_u32p = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)
_u8p = C.POINTER(C.c_uint8)


def _ptr(a, t):
    return a.ctypes.data_as(t)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class Restatement:
    """The C restatement (oracle/ttoracle.c)."""

    def __init__(self, path=None):
        path = path or os.path.join(HERE, "libttoracle.so")
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (run: make -C oracle)")
        L = self.lib = C.CDLL(path)
        L.tto_metric_parse.argtypes = [C.c_char_p, C.POINTER(C.c_int), _f32p]
        L.tto_dist.argtypes = [C.c_int, C.c_float, _f32p, _f32p, C.c_uint32]
        L.tto_dist.restype = C.c_float
        L.tto_make_dense.argtypes = [_f32p, C.c_uint32, C.c_uint32, C.c_int, C.c_float, C.c_int, _f32p]
        L.tto_enclosing_radius.argtypes = [_f32p, C.c_uint32, C.c_int]
        L.tto_enclosing_radius.restype = C.c_float
        L.tto_radius_graph.argtypes = [_f32p, C.c_uint32, C.c_uint32, C.c_int, C.c_float, C.c_float,
                                       _u64p, _u32p, _f32p]
        L.tto_radius_graph.restype = C.c_uint64
        L.tto_knn.argtypes = [_f32p, C.c_uint32, C.c_uint32, C.c_int, C.c_float, C.c_uint32, _u32p]
        L.tto_knn.restype = C.c_uint32
        L.tto_knn_rows.argtypes = [_f32p, C.c_uint32, C.c_uint32, C.c_int, C.c_float, C.c_uint32, _u32p,
                                   C.c_uint32, _u32p]
        L.tto_radius_row.argtypes = [_f32p, C.c_uint32, C.c_uint32, C.c_int, C.c_float, C.c_float, C.c_uint32,
                                     _u32p, _f32p]
        L.tto_radius_row.restype = C.c_uint64
        L.tto_knn_graph.argtypes = [_f32p, C.c_uint32, C.c_uint32, C.c_int, C.c_float, C.c_uint32,
                                    _u32p, _u64p, _u32p, _f32p]
        L.tto_knn_graph.restype = C.c_uint64
        L.tto_mst_deaths.argtypes = [_f32p, C.c_uint32, C.c_uint32, C.c_int, C.c_float, _f32p]
        L.tto_dense_edges.argtypes = [_f32p, C.c_uint32, C.c_float, _u64p, _f32p]
        L.tto_dense_edges.restype = C.c_uint64
        for fn in (L.tto_apparent_cofacet, L.tto_apparent_facet):
            fn.argtypes = [_f32p, C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint64, C.c_float]
            fn.restype = C.c_int
        L.tto_apparent_edges.argtypes = [_f32p, C.c_uint32, _u64p, _f32p, C.c_uint64, _u8p]
        L.tto_decode.argtypes = [C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint64, _u32p]
        cx = [C.c_int, C.c_uint32, _f32p, _u64p, _u32p, _f32p]
        L.tto_apparent_cx.argtypes = cx + [C.c_uint32, C.c_uint32, _u64p, _f32p, C.c_uint64, _u8p]
        L.tto_emergent_cx.argtypes = cx + [C.c_float, C.c_uint32, C.c_uint32, _u64p, _f32p, C.c_uint64, _u64p,
                                           _u8p]
        L.tto_assemble_cx.argtypes = cx + [C.c_float, C.c_uint32, C.c_uint32, _u64p, _f32p, C.c_uint64, _u64p,
                                           C.c_uint64, C.c_int, _u64p, C.POINTER(_u64p), C.POINTER(_f32p),
                                           C.POINTER(_u64p), C.POINTER(_u8p), _u64p, C.POINTER(_u64p),
                                           C.POINTER(_f32p), _u64p]
        L.tto_free.argtypes = [C.c_void_p]

    def parse(self, metric):
        k, p = C.c_int(), C.c_float()
        rc = self.lib.tto_metric_parse(metric.encode(), C.byref(k), C.byref(p))
        if rc:
            raise ValueError(f"metric: rejected '{metric}'")
        return k.value, p.value

    def dist(self, metric, a, b):
        kind, p = self.parse(metric)
        a, b = _f32(a), _f32(b)
        return self.lib.tto_dist(kind, p, _ptr(a, _f32p), _ptr(b, _f32p), a.size)

    def make_dense(self, x, metric, full=True):
        x = _f32(x)
        n, d = x.shape
        kind, p = self.parse(metric)
        out = np.zeros(n * n if full else n * (n - 1) // 2 if n else 0, np.float32)
        self.lib.tto_make_dense(_ptr(x, _f32p), n, d, kind, p, int(full), _ptr(out, _f32p))
        return out.reshape(n, n) if full else out

    def enclosing_radius(self, mat, n, full=True):
        m = _f32(mat).ravel()
        return self.lib.tto_enclosing_radius(_ptr(m, _f32p), n, int(full))

    def radius_graph(self, x, metric, tau):
        x = _f32(x)
        n, d = x.shape
        kind, p = self.parse(metric)
        rp = np.zeros(n + 1, np.uint64)
        nnz = self.lib.tto_radius_graph(_ptr(x, _f32p), n, d, kind, p, tau, _ptr(rp, _u64p), None, None)
        col = np.zeros(max(nnz, 1), np.uint32)
        w = np.zeros(max(nnz, 1), np.float32)
        self.lib.tto_radius_graph(_ptr(x, _f32p), n, d, kind, p, tau, _ptr(rp, _u64p),
                                  _ptr(col, _u32p), _ptr(w, _f32p))
        return rp, col[:nnz], w[:nnz]

    def knn(self, x, metric, k):
        x = _f32(x)
        n, d = x.shape
        kind, p = self.parse(metric)
        kk = min(k, n - 1) if n else 0
        ids = np.zeros(max(n * kk, 1), np.uint32)
        kk = self.lib.tto_knn(_ptr(x, _f32p), n, d, kind, p, k, _ptr(ids, _u32p))
        return ids[: n * kk].reshape(n, kk)

    def knn_rows(self, x, metric, k, rows):
        """knn for the given query rows only: (len(rows), k) ids, k <= n - 1."""
        x = _f32(x)
        n, d = x.shape
        kind, p = self.parse(metric)
        rows = np.ascontiguousarray(rows, np.uint32)
        ids = np.zeros(max(rows.size * k, 1), np.uint32)
        self.lib.tto_knn_rows(_ptr(x, _f32p), n, d, kind, p, k, _ptr(rows, _u32p), rows.size, _ptr(ids, _u32p))
        return ids[: rows.size * k].reshape(rows.size, k)

    def radius_row(self, x, metric, tau, i):
        """(neighbour ids, weights) of row i of the radius graph."""
        x = _f32(x)
        n, d = x.shape
        kind, p = self.parse(metric)
        m = self.lib.tto_radius_row(_ptr(x, _f32p), n, d, kind, p, tau, int(i), None, None)
        col = np.zeros(max(m, 1), np.uint32)
        w = np.zeros(max(m, 1), np.float32)
        self.lib.tto_radius_row(_ptr(x, _f32p), n, d, kind, p, tau, int(i), _ptr(col, _u32p), _ptr(w, _f32p))
        return col[:m], w[:m]

    def knn_graph(self, x, metric, ids):
        x = _f32(x)
        n, d = x.shape
        kind, p = self.parse(metric)
        ids = np.ascontiguousarray(ids, np.uint32)
        k = ids.shape[1] if ids.ndim == 2 else 0
        rp = np.zeros(n + 1, np.uint64)
        nnz = self.lib.tto_knn_graph(_ptr(x, _f32p), n, d, kind, p, k, _ptr(ids, _u32p),
                                     _ptr(rp, _u64p), None, None)
        col = np.zeros(max(nnz, 1), np.uint32)
        w = np.zeros(max(nnz, 1), np.float32)
        self.lib.tto_knn_graph(_ptr(x, _f32p), n, d, kind, p, k, _ptr(ids, _u32p), _ptr(rp, _u64p),
                               _ptr(col, _u32p), _ptr(w, _f32p))
        return rp, col[:nnz], w[:nnz]

    def mst_deaths(self, x, metric):
        x = _f32(x)
        n, d = x.shape
        kind, p = self.parse(metric)
        out = np.zeros(max(n - 1, 1), np.float32)
        self.lib.tto_mst_deaths(_ptr(x, _f32p), n, d, kind, p, _ptr(out, _f32p))
        return out[: max(n - 1, 0)]

    def dense_edges(self, mat, threshold):
        mat = _f32(mat)
        n = mat.shape[0]
        m = self.lib.tto_dense_edges(_ptr(mat, _f32p), n, threshold, None, None)
        idx = np.zeros(max(m, 1), np.uint64)
        d = np.zeros(max(m, 1), np.float32)
        self.lib.tto_dense_edges(_ptr(mat, _f32p), n, threshold, _ptr(idx, _u64p), _ptr(d, _f32p))
        return idx[:m], d[:m]

    def apparent_edges(self, mat, idx, d):
        mat = _f32(mat)
        idx = np.ascontiguousarray(idx, np.uint64)
        d = _f32(d)
        flags = np.zeros(max(idx.size, 1), np.uint8)
        self.lib.tto_apparent_edges(_ptr(mat, _f32p), mat.shape[0], _ptr(idx, _u64p), _ptr(d, _f32p),
                                    idx.size, _ptr(flags, _u8p))
        return flags[: idx.size]

    def apparent_cofacet(self, mat, nv, idx, d):
        mat = _f32(mat)
        return bool(self.lib.tto_apparent_cofacet(_ptr(mat, _f32p), mat.shape[0], nv,
                                                  idx & (2**64 - 1), idx >> 64, d))

    def apparent_facet(self, mat, nv, idx, d):
        mat = _f32(mat)
        return bool(self.lib.tto_apparent_facet(_ptr(mat, _f32p), mat.shape[0], nv,
                                                idx & (2**64 - 1), idx >> 64, d))

    def decode(self, n, nv, idx):
        vd = np.zeros(nv, np.uint32)
        self.lib.tto_decode(n, nv, idx & (2**64 - 1), idx >> 64, _ptr(vd, _u32p))
        return vd

    def apparent_cx(self, cx, nv, idx, d, words=1):
        """bit 0: zero_apparent_cofacet, bit 1: zero_apparent_facet, per simplex."""
        iw = index_words(idx, words)
        d = _f32(d)
        out = np.zeros(max(len(d), 1), np.uint8)
        self.lib.tto_apparent_cx(*cx.args(), nv, words, _ptr(iw, _u64p), _ptr(d, _f32p), len(d), _ptr(out, _u8p))
        return out[: len(d)]

    def emergent_cx(self, cx, threshold, nv, idx, d, words=1):
        """(cofacet indices, states) of compute_pairs' emergent shortcut."""
        iw = index_words(idx, words)
        d = _f32(d)
        cof = np.zeros((max(len(d), 1), words), np.uint64)
        st = np.zeros(max(len(d), 1), np.uint8)
        self.lib.tto_emergent_cx(*cx.args(), threshold, nv, words, _ptr(iw, _u64p), _ptr(d, _f32p), len(d),
                                 _ptr(cof, _u64p), _ptr(st, _u8p))
        return index_values(cof[: len(d)], words), st[: len(d)]

    def assemble_cx(self, cx, threshold, nv, idx, d, pivots, keep_next, words=1):
        """(columns [(idx, d)], emergent (cof, state) per column, next
        simplices [(idx, d)], skipped) of Reducer::assemble."""
        iw = index_words(idx, words)
        d = _f32(d)
        pv = index_words(sorted(int(v) for v in pivots), words)
        nc, nn, sk = C.c_uint64(), C.c_uint64(), C.c_uint64()
        ci, cf, nix = _u64p(), _u64p(), _u64p()
        cd, nd = _f32p(), _f32p()
        cs = _u8p()
        self.lib.tto_assemble_cx(*cx.args(), threshold, nv, words, _ptr(iw, _u64p), _ptr(d, _f32p), len(d),
                                 _ptr(pv, _u64p) if len(pv) else None, len(pv), int(keep_next), C.byref(nc),
                                 C.byref(ci), C.byref(cd), C.byref(cf), C.byref(cs), C.byref(nn), C.byref(nix),
                                 C.byref(nd), C.byref(sk))

        def take(p, k, dt):
            out = np.ctypeslib.as_array(p, shape=(max(k, 1),))[:k].astype(dt, copy=True)
            self.lib.tto_free(p)
            return out

        k, m = nc.value, nn.value
        cols = list(zip(index_values(take(ci, k * words, np.uint64), words), take(cd, k, np.float32).tolist()))
        em = (index_values(take(cf, k * words, np.uint64), words), take(cs, k, np.uint8))
        nxt = list(zip(index_values(take(nix, m * words, np.uint64), words), take(nd, m, np.float32).tolist()))
        return cols, em, nxt, sk.value


class Complex:
    """A dense (full n x n matrix) or sparse (CSR graph) complex for the
    reducer's per-simplex tests, marshalled for the C entry points."""

    def __init__(self, mat=None, n=None, row_ptr=None, col=None, w=None):
        if mat is not None:
            self.sparse = 0
            self.mat = _f32(mat)
            self.n = self.mat.shape[0]
            self.rp = self.col = self.w = None
        else:
            self.sparse = 1
            self.n = int(n)
            self.mat = None
            self.rp = np.ascontiguousarray(row_ptr, np.uint64)
            self.col = np.ascontiguousarray(col, np.uint32)
            self.w = _f32(w)

    def args(self):
        return (self.sparse, self.n,
                None if self.mat is None else _ptr(self.mat, _f32p),
                None if self.rp is None else _ptr(self.rp, _u64p),
                None if self.col is None else _ptr(self.col, _u32p),
                None if self.w is None else _ptr(self.w, _f32p))


def index_words(values, words):
    """Python ints -> (count,) or (count, 2) uint64 words, low word first."""
    vals = [int(v) for v in values]
    if words == 1:
        return np.array(vals, dtype=np.uint64)
    return np.array([[v & (2**64 - 1), v >> 64] for v in vals], dtype=np.uint64).reshape(-1, 2)


def index_values(words_arr, words):
    a = np.asarray(words_arr, dtype=np.uint64)
    if words == 1:
        return [int(v) for v in a.ravel()]
    a = a.reshape(-1, 2)
    return [int(lo) | (int(hi) << 64) for lo, hi in a]


def parse_metric(metric):
    """(kind, p) exactly as the reference's Metric::parse (metric.hpp:240-253)."""
    return Restatement().parse(metric)


def reference_available(simd=False):
    name = "libttref.so" if simd else "libttref_nosimd.so"
    return os.path.exists(os.path.join(HERE, "_ref", name))


class Reference:
    """The reference compiled from its own sources (oracle/_ref)."""

    def __init__(self, simd=False):
        name = "libttref.so" if simd else "libttref_nosimd.so"
        path = os.path.join(HERE, "_ref", name)
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (reference sources absent at build time)")
        L = self.lib = C.CDLL(path)
        L.ref_last_error.restype = C.c_char_p
        L.ref_free.argtypes = [C.c_void_p]
        L.ref_set_threads.argtypes = [C.c_uint32]
        L.ref_metric.argtypes = [C.c_char_p, _f32p, _f32p, C.c_uint32, _f32p]
        L.ref_make_dense.argtypes = [_f32p, C.c_uint32, C.c_uint32, C.c_char_p, C.c_uint64,
                                     C.POINTER(C.c_int), _f32p]
        L.ref_enclosing_radius.argtypes = [_f32p, C.c_uint32, C.c_int]
        L.ref_enclosing_radius.restype = C.c_float
        graph_out = [_u64p, _u64p, C.POINTER(_u32p), C.POINTER(_f32p)]
        L.ref_radius_graph.argtypes = [_f32p, C.c_uint32, C.c_uint32, C.c_char_p, C.c_float] + graph_out
        L.ref_knn.argtypes = [_f32p, C.c_uint32, C.c_uint32, C.c_char_p, C.c_uint32, _u32p, _u32p]
        L.ref_knn_graph.argtypes = [_f32p, C.c_uint32, C.c_uint32, C.c_char_p, C.c_uint32] + graph_out
        bars_out = [_u64p, C.POINTER(_u32p), C.POINTER(_f32p), C.POINTER(_f32p)]
        L.ref_mst_h0.argtypes = [_f32p, C.c_uint32, C.c_uint32, C.c_char_p, C.c_float] + bars_out
        L.ref_dense_edges.argtypes = [_f32p, C.c_uint32, C.c_uint32, C.c_char_p, C.c_uint64, C.c_float,
                                      _u64p, C.POINTER(_u64p), C.POINTER(_f32p), C.POINTER(_u8p), _f32p]
        L.ref_persistence.argtypes = [_f32p, C.c_uint32, C.c_uint32, C.c_char_p, C.c_uint32, C.c_float,
                                      C.c_int, C.c_uint32, C.c_int, C.c_uint64, C.c_uint32] + bars_out + [
                                      C.POINTER(C.c_double), _u64p]
        L.ref_persistence_csr.argtypes = [C.c_uint32, _u64p, _u32p, _f32p, C.c_uint32, C.c_float,
                                          C.c_int] + bars_out
        L.ref_build_edges.argtypes = [_f32p, C.c_uint32, C.c_uint32, C.c_char_p, C.c_float, _u64p,
                                      C.POINTER(C.POINTER(C.c_int32)), C.POINTER(_f32p), _f32p]
        L.ref_build_edges_csr.argtypes = [C.c_uint32, _u64p, _u32p, _f32p, C.c_float, _u64p,
                                          C.POINTER(C.POINTER(C.c_int32)), C.POINTER(_f32p)]
        L.ref_time_dense.argtypes = [_f32p, C.c_uint32, C.c_uint32, C.c_char_p, C.c_uint32,
                                     C.POINTER(C.c_double), C.POINTER(C.c_double), _f32p]
        L.ref_time_graph.argtypes = [_f32p, C.c_uint32, C.c_uint32, C.c_char_p, C.c_int, C.c_float,
                                     C.c_uint32, C.c_uint32, C.POINTER(C.c_double), _u64p]
        cx = [C.c_int, C.c_uint32, _f32p, _u64p, _u32p, _f32p]
        L.ref_apparent_cx.argtypes = cx + [C.c_uint32, C.c_uint32, _u64p, _f32p, C.c_uint64, _u8p]
        L.ref_emergent_cx.argtypes = cx + [C.c_float, C.c_uint32, C.c_uint32, _u64p, _f32p, C.c_uint64, _u64p,
                                           _u8p]
        L.ref_assemble_cx.argtypes = cx + [C.c_float, C.c_uint32, C.c_uint32, _u64p, _f32p, C.c_uint64, _u64p,
                                           C.c_uint64, C.c_int, C.c_uint32, _u64p, C.POINTER(_u64p),
                                           C.POINTER(_f32p), _u64p, C.POINTER(_u64p), C.POINTER(_f32p), _u64p]
        L.ref_compute_cx.argtypes = cx + [C.c_float, C.c_uint32, C.c_int] + bars_out + [C.POINTER(C.c_int)]

    _ERR = {1: ValueError, 2: MemoryError, 3: OverflowError, 4: RuntimeError}

    def _check(self, rc):
        if rc:
            raise self._ERR.get(rc, RuntimeError)(self.lib.ref_last_error().decode())

    def _take(self, p, n, dtype):
        out = np.ctypeslib.as_array(p, shape=(max(n, 1),))[:n].astype(dtype, copy=True)
        self.lib.ref_free(p)
        return out

    def set_threads(self, t):
        self.lib.ref_set_threads(t)

    def metric(self, metric, a, b):
        a, b = _f32(a), _f32(b)
        out = C.c_float()
        self._check(self.lib.ref_metric(metric.encode(), _ptr(a, _f32p), _ptr(b, _f32p), a.size, C.byref(out)))
        return out.value

    def make_dense(self, x, metric, budget=2**64 - 1):
        x = _f32(x)
        n, d = x.shape
        full = C.c_int()
        out = np.zeros(max(n * n, 1), np.float32)
        self._check(self.lib.ref_make_dense(_ptr(x, _f32p), n, d, metric.encode(), budget, C.byref(full),
                                            _ptr(out, _f32p)))
        if full.value:
            return True, out[: n * n].reshape(n, n)
        return False, out[: n * (n - 1) // 2]

    def enclosing_radius(self, mat, n, full=True):
        m = _f32(mat).ravel()
        return self.lib.ref_enclosing_radius(_ptr(m, _f32p), n, int(full))

    def _graph(self, fn, *args):
        n = args[1]
        rp = np.zeros(n + 1, np.uint64)
        nnz = C.c_uint64()
        col, w = _u32p(), _f32p()
        self._check(fn(*args, _ptr(rp, _u64p), C.byref(nnz), C.byref(col), C.byref(w)))
        return rp, self._take(col, nnz.value, np.uint32), self._take(w, nnz.value, np.float32)

    def radius_graph(self, x, metric, tau):
        x = _f32(x)
        n, d = x.shape
        return self._graph(self.lib.ref_radius_graph, _ptr(x, _f32p), n, d, metric.encode(), tau)

    def knn(self, x, metric, k):
        x = _f32(x)
        n, d = x.shape
        kk = C.c_uint32()
        ids = np.zeros(max(n * min(k, max(n - 1, 0)), 1), np.uint32)
        self._check(self.lib.ref_knn(_ptr(x, _f32p), n, d, metric.encode(), k, C.byref(kk), _ptr(ids, _u32p)))
        return ids[: n * kk.value].reshape(n, kk.value)

    def knn_graph(self, x, metric, k):
        x = _f32(x)
        n, d = x.shape
        return self._graph(self.lib.ref_knn_graph, _ptr(x, _f32p), n, d, metric.encode(), k)

    def _bars(self, fn, *args):
        nb = C.c_uint64()
        dims, b, e = _u32p(), _f32p(), _f32p()
        self._check(fn(*args, C.byref(nb), C.byref(dims), C.byref(b), C.byref(e)))
        k = nb.value
        return self._take(dims, k, np.uint32), self._take(b, k, np.float32), self._take(e, k, np.float32)

    def mst_h0(self, x, metric, threshold=float("inf")):
        x = _f32(x)
        n, d = x.shape
        return self._bars(self.lib.ref_mst_h0, _ptr(x, _f32p), n, d, metric.encode(), threshold)

    def dense_edges(self, x, metric, threshold=float("inf"), budget=2**64 - 1):
        x = _f32(x)
        n, d = x.shape
        m = C.c_uint64()
        idx, dd, ap = _u64p(), _f32p(), _u8p()
        thr = C.c_float()
        self._check(self.lib.ref_dense_edges(_ptr(x, _f32p), n, d, metric.encode(), budget, threshold,
                                             C.byref(m), C.byref(idx), C.byref(dd), C.byref(ap), C.byref(thr)))
        k = m.value
        return (self._take(idx, k, np.uint64), self._take(dd, k, np.float32), self._take(ap, k, np.uint8),
                thr.value)

    def build_edges(self, x, metric="l2", threshold=float("inf")):
        """redzed::detail::build_edges on the dense complex: (verts (m, 2),
        diam, threshold)."""
        x = _f32(x)
        n, d = x.shape
        m = C.c_uint64()
        pv, pd = C.POINTER(C.c_int32)(), _f32p()
        thr = C.c_float()
        self._check(self.lib.ref_build_edges(_ptr(x, _f32p), n, d, metric.encode(), threshold, C.byref(m),
                                             C.byref(pv), C.byref(pd), C.byref(thr)))
        k = m.value
        return self._take(pv, 2 * k, np.int32).reshape(k, 2), self._take(pd, k, np.float32), thr.value

    def build_edges_csr(self, n, row_ptr, col, w, threshold=float("inf")):
        rp = np.ascontiguousarray(row_ptr, np.uint64)
        col = np.ascontiguousarray(col, np.uint32)
        w = np.ascontiguousarray(w, np.float32)
        m = C.c_uint64()
        pv, pd = C.POINTER(C.c_int32)(), _f32p()
        self._check(self.lib.ref_build_edges_csr(n, _ptr(rp, _u64p), _ptr(col, _u32p), _ptr(w, _f32p), threshold,
                                                 C.byref(m), C.byref(pv), C.byref(pd)))
        k = m.value
        return self._take(pv, 2 * k, np.int32).reshape(k, 2), self._take(pd, k, np.float32)

    def persistence(self, x, maxdim=1, threshold=float("inf"), metric="l2", graph="full", k=15,
                    mode="auto", budget=0, threads=0):
        x = _f32(x)
        n, d = x.shape
        secs = C.c_double()
        edges = C.c_uint64()
        nb = C.c_uint64()
        dims, b, e = _u32p(), _f32p(), _f32p()
        g = {"full": 0, "knn": 1}[graph]
        md = {"auto": 0, "boundary": 1, "coboundary": 2}[mode]
        self._check(self.lib.ref_persistence(_ptr(x, _f32p), n, d, metric.encode(), maxdim, threshold, g, k, md,
                                             budget, threads, C.byref(nb), C.byref(dims), C.byref(b), C.byref(e),
                                             C.byref(secs), C.byref(edges)))
        kk = nb.value
        return ((self._take(dims, kk, np.uint32), self._take(b, kk, np.float32), self._take(e, kk, np.float32)),
                {"seconds_total": secs.value, "edges": edges.value})

    def persistence_csr(self, n, row_ptr, col, w, maxdim=1, threshold=float("inf"), mode="auto"):
        rp = np.ascontiguousarray(row_ptr, np.uint64)
        col = np.ascontiguousarray(col, np.uint32)
        w = _f32(w)
        md = {"auto": 0, "boundary": 1, "coboundary": 2}[mode]
        return self._bars(self.lib.ref_persistence_csr, n, _ptr(rp, _u64p), _ptr(col, _u32p), _ptr(w, _f32p),
                          maxdim, threshold, md)

    def apparent_cx(self, cx, nv, idx, d, words=1):
        iw = index_words(idx, words)
        d = _f32(d)
        out = np.zeros(max(len(d), 1), np.uint8)
        self._check(self.lib.ref_apparent_cx(*cx.args(), nv, words, _ptr(iw, _u64p), _ptr(d, _f32p), len(d),
                                             _ptr(out, _u8p)))
        return out[: len(d)]

    def emergent_cx(self, cx, threshold, nv, idx, d, words=1):
        iw = index_words(idx, words)
        d = _f32(d)
        cof = np.zeros((max(len(d), 1), words), np.uint64)
        st = np.zeros(max(len(d), 1), np.uint8)
        self._check(self.lib.ref_emergent_cx(*cx.args(), threshold, nv, words, _ptr(iw, _u64p), _ptr(d, _f32p),
                                             len(d), _ptr(cof, _u64p), _ptr(st, _u8p)))
        return index_values(cof[: len(d)], words), st[: len(d)]

    def assemble_cx(self, cx, threshold, nv, idx, d, pivots, keep_next, words=1, threads=1):
        """(columns [(idx, d)], next simplices [(idx, d)], skipped) of
        Reducer::assemble."""
        iw = index_words(idx, words)
        d = _f32(d)
        pv = index_words(sorted(int(v) for v in pivots), words)
        nc, nn, sk = C.c_uint64(), C.c_uint64(), C.c_uint64()
        ci, nix = _u64p(), _u64p()
        cd, nd = _f32p(), _f32p()
        self._check(self.lib.ref_assemble_cx(*cx.args(), threshold, nv, words, _ptr(iw, _u64p), _ptr(d, _f32p),
                                             len(d), _ptr(pv, _u64p) if len(pv) else None, len(pv), int(keep_next),
                                             threads, C.byref(nc), C.byref(ci), C.byref(cd), C.byref(nn),
                                             C.byref(nix), C.byref(nd), C.byref(sk)))
        k, m = nc.value, nn.value
        cols = list(zip(index_values(self._take(ci, k * words, np.uint64), words),
                        self._take(cd, k, np.float32).tolist()))
        nxt = list(zip(index_values(self._take(nix, m * words, np.uint64), words),
                       self._take(nd, m, np.float32).tolist()))
        return cols, nxt, sk.value

    def compute_cx(self, cx, threshold, maxdim, force_wide=False):
        """rips::compute with force_wide: ((dims, births, deaths), wide)."""
        nb = C.c_uint64()
        dims, b, e = _u32p(), _f32p(), _f32p()
        wide = C.c_int()
        self._check(self.lib.ref_compute_cx(*cx.args(), threshold, maxdim, int(force_wide), C.byref(nb),
                                            C.byref(dims), C.byref(b), C.byref(e), C.byref(wide)))
        k = nb.value
        return ((self._take(dims, k, np.uint32), self._take(b, k, np.float32), self._take(e, k, np.float32)),
                bool(wide.value))

    def time_dense(self, x, metric="l2", threads=0):
        x = _f32(x)
        n, d = x.shape
        a, b = C.c_double(), C.c_double()
        r = C.c_float()
        self._check(self.lib.ref_time_dense(_ptr(x, _f32p), n, d, metric.encode(), threads, C.byref(a),
                                            C.byref(b), C.byref(r)))
        return a.value, b.value, r.value

    def time_graph(self, x, metric, which, tau=float("inf"), k=15, threads=0):
        x = _f32(x)
        n, d = x.shape
        s = C.c_double()
        size = C.c_uint64()
        w = {"radius": 0, "knn": 1, "mst": 2}[which]
        self._check(self.lib.ref_time_graph(_ptr(x, _f32p), n, d, metric.encode(), w, tau, k, threads,
                                            C.byref(s), C.byref(size)))
        return s.value, size.value
