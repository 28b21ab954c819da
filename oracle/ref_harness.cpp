// SPDX-License-Identifier: MIT
// ref_harness.cpp — TEST INFRASTRUCTURE ONLY (oracle side).
//
// A C ABI over the reference library's own front-end functions, compiled
// against the reference headers where they lie (see oracle/Makefile) into
// oracle/_ref/. Nothing here is shipped or measured as the product: the
// parity tests call it as the checker, and bench.py times it as the CPU
// baseline. Every entry point names the reference function it drives.
//
// Variable-size outputs are returned in malloc'd buffers released by
// ref_free().
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <memory>
#include <type_traits>
#include <string>

#include "turbotops/turbotops.hpp"

using namespace tops;

namespace {
thread_local std::string g_err;

template <typename F>
int guard(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::invalid_argument& e) {
        g_err = e.what();
        return 1;
    } catch (const std::length_error& e) {
        g_err = e.what();
        return 2;
    } catch (const std::overflow_error& e) {
        g_err = e.what();
        return 3;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 4;
    }
}

template <typename T>
T* dup(const std::vector<T>& v) {
    T* p = static_cast<T*>(std::malloc(std::max<size_t>(1, v.size() * sizeof(T))));
    if (!v.empty()) std::memcpy(p, v.data(), v.size() * sizeof(T));
    return p;
}

void flatten(const rips::Barcode& bc, uint64_t* nbars, uint32_t** dims, float** births,
             float** deaths) {
    std::vector<uint32_t> d;
    std::vector<float> b, e;
    for (u32 k = 0; k < bc.ndims(); ++k)
        for (u64 i = 0; i < bc.birth[k].size(); ++i) {
            d.push_back(k);
            b.push_back(bc.birth[k][i]);
            e.push_back(bc.death[k][i]);
        }
    *nbars = d.size();
    *dims = dup(d);
    *births = dup(b);
    *deaths = dup(e);
}

double secs_since(std::chrono::steady_clock::time_point t) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t).count();
}
}  // namespace

extern "C" {

const char* ref_last_error() { return g_err.c_str(); }
void ref_free(void* p) { std::free(p); }
void ref_set_threads(uint32_t t) { set_num_threads(t); }

// kern::l1/l2/linf/lp through Metric (include/turbotops/metric.hpp:92-225).
int ref_metric(const char* metric, const float* a, const float* b, uint32_t dim, float* out) {
    return guard([&] { *out = Metric::parse(metric)(a, b, dim); });
}

// rips::make_dense with a PointCloud functor (rips.hpp:391-416). *full
// reports the chosen layout; out must hold n*n floats (full) or n(n-1)/2.
int ref_make_dense(const float* pts, uint32_t n, uint32_t dim, const char* metric,
                   uint64_t budget, int* full, float* out) {
    return guard([&] {
        const PointCloud c{pts, n, dim, Metric::parse(metric)};
        rips::DenseComplex cx = rips::make_dense(n, c, budget);
        *full = cx.full ? 1 : 0;
        if (!cx.mat.empty()) std::memcpy(out, cx.mat.data(), cx.mat.size() * sizeof(float));
    });
}

// DenseComplex::enclosing_radius over a matrix in either layout (rips.hpp:370-380).
float ref_enclosing_radius(const float* mat, uint32_t n, int full) {
    rips::DenseComplex cx;
    cx.n = n;
    cx.full = full != 0;
    cx.mat.assign(mat, mat + (full ? u64(n) * n : rips::dense_bytes(n, false) / 4));
    return cx.enclosing_radius();
}

// radius_graph (graph.hpp:170-201).
int ref_radius_graph(const float* pts, uint32_t n, uint32_t dim, const char* metric, float tau,
                     uint64_t* row_ptr, uint64_t* nnz, uint32_t** col, float** w) {
    return guard([&] {
        const PointCloud c{pts, n, dim, Metric::parse(metric)};
        const CsrGraph g = radius_graph(c, tau);
        std::memcpy(row_ptr, g.row_ptr.data(), (u64(n) + 1) * sizeof(u64));
        *nnz = g.col_idx.size();
        *col = dup(g.col_idx);
        *w = dup(g.weights);
    });
}

// knn_lists_exact (graph.hpp:211-225); ids holds n * min(k, n-1).
int ref_knn(const float* pts, uint32_t n, uint32_t dim, const char* metric, uint32_t k,
            uint32_t* kk, uint32_t* ids) {
    return guard([&] {
        const PointCloud c{pts, n, dim, Metric::parse(metric)};
        const KnnLists L = knn_lists_exact(c, k);
        *kk = L.k;
        if (!L.ids.empty()) std::memcpy(ids, L.ids.data(), L.ids.size() * sizeof(u32));
    });
}

// knn_graph of exact lists (graph.hpp:286-304).
int ref_knn_graph(const float* pts, uint32_t n, uint32_t dim, const char* metric, uint32_t k,
                  uint64_t* row_ptr, uint64_t* nnz, uint32_t** col, float** w) {
    return guard([&] {
        const PointCloud c{pts, n, dim, Metric::parse(metric)};
        const CsrGraph g = knn_graph(c, knn_lists_exact(c, k));
        std::memcpy(row_ptr, g.row_ptr.data(), (u64(n) + 1) * sizeof(u64));
        *nnz = g.col_idx.size();
        *col = dup(g.col_idx);
        *w = dup(g.weights);
    });
}

// mst::h0 (mst.hpp:25-99): bars flattened in discovery order.
int ref_mst_h0(const float* pts, uint32_t n, uint32_t dim, const char* metric, float threshold,
               uint64_t* nbars, uint32_t** dims, float** births, float** deaths) {
    return guard([&] {
        const PointCloud c{pts, n, dim, Metric::parse(metric)};
        flatten(mst::h0(c, threshold), nbars, dims, births, deaths);
    });
}

// The coboundary reducer's dim-0 inputs on the dense complex at the
// enclosing radius (rips.hpp:685-849): the kept edges sorted by `precedes`
// (diameter ascending, index descending) and, per edge, whether
// zero_apparent_cofacet finds its apparent partner.
int ref_dense_edges(const float* pts, uint32_t n, uint32_t dim, const char* metric,
                    uint64_t budget, float threshold, uint64_t* m, uint64_t** idx, float** d,
                    uint8_t** apparent, float* thr_out) {
    return guard([&] {
        const PointCloud c{pts, n, dim, Metric::parse(metric)};
        rips::DenseComplex cx = rips::make_dense(n, c, budget);
        cx.threshold = std::isinf(threshold) ? cx.enclosing_radius() : threshold;
        *thr_out = cx.threshold;
        using E = rips::Entry<u64>;
        std::vector<E> edges;
        for (u32 i = 1; i < n; ++i)
            for (u32 j = 0; j < i; ++j) {
                const f32 dd = cx.dist(i, j);
                if (dd <= cx.threshold) edges.push_back(E{u64(i) * (i - 1) / 2 + j, dd});
            }
        std::sort(edges.begin(), edges.end(),
                  [](const E& a, const E& b) { return rips::precedes(a, b); });
        rips::Barcode out;
        rips::RipsStats st;
        rips::Reducer<u64, rips::DenseComplex> r(cx, 1, out, st);
        std::vector<u64> ix(edges.size());
        std::vector<float> dv(edges.size());
        std::vector<uint8_t> ap(edges.size());
        for (size_t t = 0; t < edges.size(); ++t) {
            ix[t] = edges[t].idx;
            dv[t] = edges[t].d;
            ap[t] = r.zero_apparent_cofacet(edges[t], 2).idx != rips::kNone<u64>;
        }
        *m = edges.size();
        *idx = dup(ix);
        *d = dup(dv);
        *apparent = dup(ap);
    });
}

// redzed::detail::build_edges (redzed.hpp:544-613) over the dense complex of
// the points at `threshold` (+inf: the enclosing radius, as
// turbotops.hpp:245 sets it): EdgeList's verts (2 per edge) and diam.
int ref_build_edges(const float* pts, uint32_t n, uint32_t dim, const char* metric, float threshold,
                    uint64_t* m, int32_t** verts, float** diam, float* thr_out) {
    return guard([&] {
        const PointCloud c{pts, n, dim, Metric::parse(metric)};
        rips::DenseComplex cx = rips::make_dense(n, c, ~u64(0));
        cx.threshold = std::isinf(threshold) ? cx.enclosing_radius() : threshold;
        *thr_out = cx.threshold;
        redzed::detail::DenseGeom geo(cx);
        redzed::detail::EdgeList el;
        const redzed::detail::BinomialTable bt(int(n), 3);
        redzed::detail::build_edges(geo, cx.threshold, bt, el);
        *m = el.count;
        el.verts.resize(2 * el.count);
        el.diam.resize(el.count);
        *verts = dup(el.verts);
        *diam = dup(el.diam);
    });
}

// build_edges over a CSR graph (SparseGeom, redzed.hpp:831-979).
int ref_build_edges_csr(uint32_t n, const uint64_t* row_ptr, const uint32_t* col, const float* w,
                        float threshold, uint64_t* m, int32_t** verts, float** diam) {
    return guard([&] {
        CsrGraph g;
        g.n = n;
        g.row_ptr.assign(row_ptr, row_ptr + n + 1);
        g.col_idx.assign(col, col + row_ptr[n]);
        g.weights.assign(w, w + row_ptr[n]);
        g.m = row_ptr[n] / 2;
        redzed::detail::SparseGeom<uint32_t> geo(g);
        redzed::detail::EdgeList el;
        const redzed::detail::BinomialTable bt(int(n), 3);
        redzed::detail::build_edges(geo, threshold, bt, el);
        *m = el.count;
        el.verts.resize(2 * el.count);
        el.diam.resize(el.count);
        *verts = dup(el.verts);
        *diam = dup(el.diam);
    });
}

// tops::persistence on points (turbotops.hpp:211-303). graph: 0 full, 1 knn;
// mode: 0 auto, 1 boundary, 2 coboundary.
int ref_persistence(const float* pts, uint32_t n, uint32_t dim, const char* metric,
                    uint32_t maxdim, float threshold, int graph, uint32_t k, int mode,
                    uint64_t budget, uint32_t threads, uint64_t* nbars, uint32_t** dims,
                    float** births, float** deaths, double* seconds_total, uint64_t* edges) {
    return guard([&] {
        Options o;
        o.maxdim = maxdim;
        o.threshold = threshold;
        o.metric = metric;
        o.graph = graph ? Options::Graph::Knn : Options::Graph::Full;
        o.k = k;
        o.mode = mode == 1 ? Mode::Boundary : mode == 2 ? Mode::Coboundary : Mode::Auto;
        o.memory_budget = budget;
        o.threads = threads;
        const Result r = persistence(pts, n, dim, o);
        flatten(r.barcode, nbars, dims, births, deaths);
        *seconds_total = r.stats.seconds_total;
        *edges = r.stats.edges;
    });
}

// tops::persistence on a CsrGraph (turbotops.hpp:179-209).
int ref_persistence_csr(uint32_t n, const uint64_t* row_ptr, const uint32_t* col,
                        const float* w, uint32_t maxdim, float threshold, int mode,
                        uint64_t* nbars, uint32_t** dims, float** births, float** deaths) {
    return guard([&] {
        CsrGraph g;
        g.n = n;
        g.row_ptr.assign(row_ptr, row_ptr + u64(n) + 1);
        g.m = g.row_ptr[n] / 2;
        g.col_idx.assign(col, col + g.row_ptr[n]);
        g.weights.assign(w, w + g.row_ptr[n]);
        const Result r = persistence(g, maxdim, threshold, 0,
                                     mode == 1 ? Mode::Boundary
                                     : mode == 2 ? Mode::Coboundary : Mode::Auto);
        flatten(r.barcode, nbars, dims, births, deaths);
    });
}

// Stage timer for the CPU baseline: make_dense then enclosing_radius, with
// steady_clock around each (the reference's own threaded code).
int ref_time_dense(const float* pts, uint32_t n, uint32_t dim, const char* metric,
                   uint32_t threads, double* s_make, double* s_radius, float* radius) {
    return guard([&] {
        if (threads) set_num_threads(threads);
        const PointCloud c{pts, n, dim, Metric::parse(metric)};
        const auto t0 = std::chrono::steady_clock::now();
        rips::DenseComplex cx = rips::make_dense(n, c);
        *s_make = secs_since(t0);
        const auto t1 = std::chrono::steady_clock::now();
        *radius = cx.enclosing_radius();
        *s_radius = secs_since(t1);
    });
}

// Stage timers for the sparse front ends: radius_graph, knn_lists_exact +
// knn_graph, mst::h0.
int ref_time_graph(const float* pts, uint32_t n, uint32_t dim, const char* metric, int which,
                   float tau, uint32_t k, uint32_t threads, double* secs, uint64_t* size) {
    return guard([&] {
        if (threads) set_num_threads(threads);
        const PointCloud c{pts, n, dim, Metric::parse(metric)};
        const auto t0 = std::chrono::steady_clock::now();
        if (which == 0) {
            *size = radius_graph(c, tau).m;
        } else if (which == 1) {
            *size = knn_graph(c, knn_lists_exact(c, k)).m;
        } else {
            *size = mst::h0(c, tau).birth[0].size();
        }
        *secs = secs_since(t0);
    });
}

}  // extern "C"

// ---- the reducer's per-simplex tests (rips.hpp:585-665, :903-928, :981-1023)
// Complexes: sparse = 0 -> DenseComplex (full layout, n x n matrix), 1 ->
// SparseComplex over a CSR graph. Simplex indices as `words` 64-bit words
// (1: Reducer<u64>, 2: Reducer<u128>, low word first).
namespace {

template <typename I>
I get_idx(const uint64_t* w, uint32_t words, uint64_t t) {
    if constexpr (sizeof(I) > 8) return words == 2 ? (I(w[2 * t + 1]) << 64) | w[2 * t] : I(w[t]);
    else return I(w[t]);
}

template <typename I>
void put_idx(uint64_t* w, uint32_t words, uint64_t t, I v) {
    if (words == 2) {
        w[2 * t] = uint64_t(v);
        if constexpr (sizeof(I) > 8) w[2 * t + 1] = uint64_t(v >> 64);
        else w[2 * t + 1] = 0;
    } else {
        w[t] = uint64_t(v);
    }
}

struct Cx {
    rips::DenseComplex dense;
    CsrGraph g;
    std::unique_ptr<rips::SparseComplex> sp;
    Cx(int sparse, uint32_t n, const float* mat, const uint64_t* rp, const uint32_t* col, const float* w,
       float threshold) {
        if (!sparse) {
            dense.n = n;
            dense.full = true;
            dense.mat.assign(mat, mat + u64(n) * n);
            dense.threshold = threshold;
        } else {
            g.n = n;
            g.row_ptr.assign(rp, rp + u64(n) + 1);
            g.m = g.row_ptr[n] / 2;
            g.col_idx.assign(col, col + g.row_ptr[n]);
            g.weights.assign(w, w + g.row_ptr[n]);
            sp = std::make_unique<rips::SparseComplex>(g, threshold);
        }
    }
};

template <typename F>
void with_reducer(const Cx& cx, uint32_t words, uint32_t maxdim, F&& f) {
    rips::Barcode out;
    rips::RipsStats st;
    if (cx.sp) {
        if (words == 2) {
            rips::Reducer<rips::u128, rips::SparseComplex> r(*cx.sp, maxdim, out, st);
            f(r, st);
        } else {
            rips::Reducer<u64, rips::SparseComplex> r(*cx.sp, maxdim, out, st);
            f(r, st);
        }
    } else {
        if (words == 2) {
            rips::Reducer<rips::u128, rips::DenseComplex> r(cx.dense, maxdim, out, st);
            f(r, st);
        } else {
            rips::Reducer<u64, rips::DenseComplex> r(cx.dense, maxdim, out, st);
            f(r, st);
        }
    }
}

}  // namespace

extern "C" {

// Reducer::zero_apparent_cofacet / zero_apparent_facet (rips.hpp:644-654) of
// simplices with nv vertices: bit 0 cofacet, bit 1 facet.
int ref_apparent_cx(int sparse, uint32_t n, const float* mat, const uint64_t* rp, const uint32_t* col,
                    const float* w, uint32_t nv, uint32_t words, const uint64_t* idx, const float* d,
                    uint64_t count, uint8_t* out) {
    return guard([&] {
        const Cx cx(sparse, n, mat, rp, col, w, INFINITY);
        with_reducer(cx, words, nv, [&](auto& r, rips::RipsStats&) {
            using R = std::remove_reference_t<decltype(r)>;
            using I = decltype(R::E::idx);
            for (uint64_t t = 0; t < count; ++t) {
                const typename R::E e{get_idx<I>(idx, words, t), d[t]};
                out[t] = uint8_t((r.zero_apparent_cofacet(e, nv).idx != rips::kNone<I> ? 1 : 0) |
                                 (r.zero_apparent_facet(e, nv).idx != rips::kNone<I> ? 2 : 0));
            }
        });
    });
}

// compute_pairs' emergent shortcut without the pivot lookup
// (rips.hpp:903-928): the first cofacet of diameter col.d in the
// enumeration order (descending index, cap = threshold) -> cof, state 1,
// or 2 when zero_apparent_facet finds that cofacet's apparent facet; 0 when
// there is none.
int ref_emergent_cx(int sparse, uint32_t n, const float* mat, const uint64_t* rp, const uint32_t* col,
                    const float* w, float threshold, uint32_t nv, uint32_t words, const uint64_t* idx,
                    const float* d, uint64_t count, uint64_t* cof, uint8_t* state) {
    return guard([&] {
        const Cx cx(sparse, n, mat, rp, col, w, threshold);
        with_reducer(cx, words, nv, [&](auto& r, rips::RipsStats&) {
            using R = std::remove_reference_t<decltype(r)>;
            using E = typename R::E;
            using I = decltype(E::idx);
            u32 vd[R::KMAX];
            for (uint64_t t = 0; t < count; ++t) {
                const E col_e{get_idx<I>(idx, words, t), d[t]};
                r.decode(col_e.idx, nv, vd);
                E cand{rips::kNone<I>, 0.0f};
                r.cofacet_entries(vd, nv, col_e.d, true, r.cx.threshold, [&](const E& cf) {
                    if (cf.d == col_e.d) {
                        cand = cf;
                        return false;
                    }
                    return true;
                });
                if (cand.idx == rips::kNone<I>) {
                    state[t] = 0;
                    put_idx<I>(cof, words, t, I(0));
                } else {
                    state[t] = r.zero_apparent_facet(cand, nv + 1).idx == rips::kNone<I> ? 1 : 2;
                    put_idx<I>(cof, words, t, cand.idx);
                }
            }
        });
    });
}

// Reducer::assemble(nv - 1) (rips.hpp:981-1023) on the given simplices (nv
// vertices) with the given pivots cleared; columns in its (sorted) order,
// the next simplices in its order when keep_next, the skipped count from
// RipsStats::apparent. Outputs malloc'd (ref_free).
int ref_assemble_cx(int sparse, uint32_t n, const float* mat, const uint64_t* rp, const uint32_t* col,
                    const float* w, float threshold, uint32_t nv, uint32_t words, const uint64_t* idx,
                    const float* d, uint64_t count, const uint64_t* pivots, uint64_t npivots, int keep_next,
                    uint32_t threads, uint64_t* ncols, uint64_t** cidx, float** cd, uint64_t* nnext,
                    uint64_t** nidx, float** nd, uint64_t* skipped) {
    return guard([&] {
        if (threads) set_num_threads(threads);
        const Cx cx(sparse, n, mat, rp, col, w, threshold);
        const u32 dim = nv - 1;
        with_reducer(cx, words, keep_next ? dim + 2 : dim + 1, [&](auto& r, rips::RipsStats& st) {
            using R = std::remove_reference_t<decltype(r)>;
            using E = typename R::E;
            using I = decltype(E::idx);
            r.simplices.clear();
            for (uint64_t t = 0; t < count; ++t) r.simplices.push_back(E{get_idx<I>(idx, words, t), d[t]});
            r.pivots.reset(2 * npivots + 16);
            for (uint64_t t = 0; t < npivots; ++t) r.pivots.insert(get_idx<I>(pivots, words, t), t);
            r.assemble(dim);
            *ncols = r.columns.size();
            *cidx = static_cast<uint64_t*>(std::malloc(std::max<size_t>(1, r.columns.size()) * words * 8));
            *cd = static_cast<float*>(std::malloc(std::max<size_t>(1, r.columns.size()) * 4));
            for (size_t t = 0; t < r.columns.size(); ++t) {
                put_idx<I>(*cidx, words, t, r.columns[t].idx);
                (*cd)[t] = r.columns[t].d;
            }
            *nnext = r.simplices.size();
            *nidx = static_cast<uint64_t*>(std::malloc(std::max<size_t>(1, r.simplices.size()) * words * 8));
            *nd = static_cast<float*>(std::malloc(std::max<size_t>(1, r.simplices.size()) * 4));
            for (size_t t = 0; t < r.simplices.size(); ++t) {
                put_idx<I>(*nidx, words, t, r.simplices[t].idx);
                (*nd)[t] = r.simplices[t].d;
            }
            *skipped = st.apparent.empty() ? 0 : st.apparent.back();
        });
    });
}

// tops::rips::compute on a dense matrix or a CSR graph with force_wide
// (rips.hpp:1038-1061): 128-bit indices even where 64 bits suffice.
int ref_compute_cx(int sparse, uint32_t n, const float* mat, const uint64_t* rp, const uint32_t* col,
                   const float* w, float threshold, uint32_t maxdim, int force_wide, uint64_t* nbars,
                   uint32_t** dims, float** births, float** deaths, int* wide) {
    return guard([&] {
        const Cx cx(sparse, n, mat, rp, col, w, threshold);
        rips::RipsStats st;
        const rips::Barcode bc = sparse ? rips::compute(*cx.sp, maxdim, &st, force_wide != 0)
                                        : rips::compute(cx.dense, maxdim, &st, force_wide != 0);
        flatten(bc, nbars, dims, births, deaths);
        *wide = st.wide_index ? 1 : 0;
    });
}

}  // extern "C"
