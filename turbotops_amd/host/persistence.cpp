// SPDX-License-Identifier: MIT
// persistence.cpp — tops::persistence and the front-end functions of the
// reference's API, with every data-parallel stage on the GPU (include/ttgpu.h)
// and the residual reduction on the host (cohomology.hpp, zerodiff.hpp).
#include <sys/resource.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <thread>

#include "turbotops_gpu.hpp"
#include "zerodiff.hpp"

namespace tops {

namespace {

std::atomic<u32> g_threads{0};

double since(std::chrono::steady_clock::time_point t) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t).count();
}

template <class T>
std::vector<T> take(T* p, u64 count) {
    std::vector<T> v;
    if (p) {
        v.assign(p, p + count);
        tt_free(p);
    }
    return v;
}

// validate_options (turbotops.hpp:170-175)
void validate_options(f32 threshold, u32 k) {
    if (std::isnan(threshold) || threshold < 0.0f)
        throw std::invalid_argument("persistence: threshold must be non-negative or +inf");
    if (k == 0) throw std::invalid_argument("persistence: k must be >= 1");
}

// redzed::fits (redzed.hpp:1432-1441): 64-bit simplex indices up to maxdim + 2.
bool boundary_fits(u32 n, u32 maxdim) {
    try {
        cob::Binomials<u64> b(std::min<u32>(n, 0x7fffffffu), maxdim + 2);
        return true;
    } catch (const std::overflow_error&) {
        return false;
    }
}

// The boundary engine (zerodiff.hpp) serves what the reference's auto mode
// sends to RedZeD on dense input (turbotops.hpp:144-152): H1 (maxdim = 1) of
// a full Rips complex; --mode boundary with maxdim 1 as well. Everything else
// runs on the cohomology engine, which gives the same diagram.
bool boundary_engine(Mode mode, u32 n, u32 maxdim, bool dense_full) {
    return dense_full && maxdim == 1 && mode != Mode::Coboundary && boundary_fits(n, maxdim);
}

// The reference's select_mode (turbotops.hpp:121-162): its checks and errors,
// and the engine that serves each case here.
std::string select_mode(Mode mode, u32 n, u32 maxdim, double mean_degree) {
    const bool dense = std::isinf(mean_degree);
    if (mode == Mode::Coboundary)
        return "coboundary (--mode coboundary): cohomology reducer, maxdim <= 30, 64/128-bit indices, any input";
    if (mode == Mode::Boundary) {
        if (maxdim == 0)
            return "boundary (--mode boundary) requested, but H0 runs as a minimum spanning tree and uses no "
                   "reduction engine";
        if (!boundary_fits(n, maxdim))
            throw std::invalid_argument(
                "boundary mode (RedZeD) needs every simplex index in 64 bits (C(n, maxdim + 2) < 2^64, "
                "implied n < 2^31); this problem does not fit — use the default coboundary mode instead");
        if (maxdim + 2 > 32) throw std::invalid_argument("redzed: maxdim must be <= 30");
        if (boundary_engine(mode, n, maxdim, dense))
            return "boundary (--mode boundary): reduction to zero differentials on the GPU's edge order";
        return "boundary (--mode boundary): served by the cohomology reducer (identical diagram)";
    }
    if (maxdim == 0) return "auto: H0 only, minimum spanning tree (no reduction engine)";
    if (boundary_engine(mode, n, maxdim, dense))
        return "auto: boundary (reduction to zero differentials) for H1 of dense full-Rips input";
    return "auto: cohomology reducer on GPU front-end outputs";
}

// H0 and H1 by the boundary engine: full matrix (row stride n) and every
// edge up to the threshold in (diameter, combinatorial index) order.
// witness: bit 31 of each hi id carries the ready-witness flag.
void reduce_boundary(u32 n, const f32* mat, u64 m, const u32* verts, const f32* d, bool witness, Barcode& out,
                     Stats& st) {
    const auto t0 = std::chrono::steady_clock::now();
    out.ensure(1);
    if (n) {
        zd::BoundaryH1 eng(n, mat, n);
        eng.run(m, verts, d, out, witness);
    }
    st.edges = m;
    st.seconds_reduce = since(t0);
}

// The GPU's column stages (include/ttgpu.h tt_complex_*) as the cohomology
// engine's ColumnSource: the complex is copied to the device once, the
// simplices between dimensions stay there.
class GpuColumns final : public cob::ColumnSource {
  public:
    GpuColumns(const cob::DenseView& cx, u32 maxdim, int words) {
        check(tt_complex_dense(cx.mat, cx.n, cx.full ? TT_LAYOUT_FULL : TT_LAYOUT_LOWER, 0, cx.threshold, maxdim,
                               words, &h_));
    }
    GpuColumns(const cob::SparseView& cx, u32 maxdim, int words) {
        check(tt_complex_sparse(cx.n, cx.row_ptr, cx.col, cx.w, cx.threshold, maxdim, words, &h_));
    }
    ~GpuColumns() override { tt_complex_destroy(h_); }
    GpuColumns(const GpuColumns&) = delete;
    GpuColumns& operator=(const GpuColumns&) = delete;

    void emergent(u32 nv, const u64* idx, const f32* d, u64 count, u64* cof, u8* state) override {
        check(tt_complex_emergent(h_, nv, idx, d, count, cof, state));
    }
    void seed(u32 nv, const u64* idx, const f32* d, u64 count) override {
        check(tt_complex_seed(h_, nv, idx, d, count));
    }
    void assemble(const u64* pivots, u64 npivots, bool keep_next, std::vector<u64>& idx, std::vector<f32>& d,
                  std::vector<u64>& cof, std::vector<u8>& state, u64& skipped) override {
        u64 nc = 0;
        u64 *pi = nullptr, *pc = nullptr;
        f32* pd = nullptr;
        u8* ps = nullptr;
        const tt_status rc = tt_complex_assemble(h_, pivots, npivots, keep_next ? 1 : 0, &nc, &pi, &pd, &pc, &ps,
                                                 &skipped);
        std::unique_ptr<u64, void (*)(void*)> a(pi, tt_free), c(pc, tt_free);
        std::unique_ptr<f32, void (*)(void*)> b(pd, tt_free);
        std::unique_ptr<u8, void (*)(void*)> e(ps, tt_free);
        check(rc);
        const u64 w = words_;
        idx.assign(pi, pi + nc * w);
        cof.assign(pc, pc + nc * w);
        d.assign(pd, pd + nc);
        state.assign(ps, ps + nc);
    }
    void set_words(u64 w) { words_ = w; }

  private:
    tt_complex* h_ = nullptr;
    u64 words_ = 1;
};

// Runs the cohomology engine on the given dimension-0 edges: dim0 on the
// host, every higher dimension's columns from the GPU.
template <class Cx>
void reduce_all(const Cx& cx, u32 maxdim, const u64* idx, const f32* d, const u8* apparent, u64 m,
                Barcode& out, Stats& st, bool force_wide = false) {
    const auto t0 = std::chrono::steady_clock::now();
    if (maxdim + 2 > 32) throw std::invalid_argument("rips: maxdim must be <= 30");
    out.ensure(maxdim);
    if (cx.n == 0) return;
    cob::Telemetry tel;
    auto run = [&]<class I>() {
        cob::Cohomology<I, Cx> r(cx, maxdim, out, tel);
        r.dim0(idx, d, apparent, m);
        if (maxdim >= 1) {
            GpuColumns src(cx, maxdim, int(sizeof(I) / 8));
            src.set_words(sizeof(I) / 8);
            r.run(src);
        }
    };
    if (!force_wide && cob::index_fits<u64>(cx.n, maxdim)) {
        run.template operator()<u64>();
    } else if (cob::index_fits<cob::u128>(cx.n, maxdim)) {
        tel.wide_index = true;
        run.template operator()<cob::u128>();
    } else {
        throw std::overflow_error("rips: C(n, maxdim + 2) exceeds 128-bit simplex indices; lower maxdim");
    }
    st.edges = tel.edges;
    st.wide_index = tel.wide_index;
    st.columns = tel.columns;
    st.apparent = tel.apparent;
    st.emergent = tel.emergent;
    st.seconds_reduce = since(t0);
}

// dim0 inputs of a sparse complex, from the GPU (tt_sparse_edges): kept
// edges (u < v, w <= threshold) in filtration order and their apparent flags.
struct SparseEdgeList {
    u64 m = 0;
    u64* idx = nullptr;
    f32* d = nullptr;
    u8* apparent = nullptr;
    SparseEdgeList(const CsrGraph& g, f32 threshold, bool want_apparent) {
        check(tt_sparse_edges(g.n, g.row_ptr.data(), g.col_idx.data(), g.weights.data(), threshold, &m, &idx,
                              &d, want_apparent ? &apparent : nullptr));
    }
    ~SparseEdgeList() {
        tt_free(idx);
        tt_free(d);
        tt_free(apparent);
    }
    SparseEdgeList(const SparseEdgeList&) = delete;
    SparseEdgeList& operator=(const SparseEdgeList&) = delete;
};

}  // namespace

// ---- globals --------------------------------------------------------------
void set_num_threads(u32 t) { g_threads.store(t); }
u32 num_threads() {
    const u32 t = g_threads.load();
    return t ? t : std::max<u32>(1, std::thread::hardware_concurrency());
}
u64 default_memory_budget() {
    const long pages = sysconf(_SC_PHYS_PAGES), page = sysconf(_SC_PAGE_SIZE);
    const u64 m = pages > 0 && page > 0 ? u64(pages) * u64(page) : 0;
    return m ? m / 2 : (u64(8) << 30);
}
double peak_rss_mb() {
    struct rusage ru;
    getrusage(RUSAGE_SELF, &ru);
    return double(ru.ru_maxrss) / 1024.0;
}

Mode parse_mode(const std::string& s) {
    if (s == "auto") return Mode::Auto;
    if (s == "boundary") return Mode::Boundary;
    if (s == "coboundary") return Mode::Coboundary;
    throw std::invalid_argument("mode must be auto, boundary or coboundary");
}

// ---- CsrGraph (core.hpp:92-149) ---------------------------------------------
f32 CsrGraph::weight(u32 u, u32 v) const {
    const u32* b = col_idx.data() + begin(u);
    const u32* e = col_idx.data() + end(u);
    const u32* it = std::lower_bound(b, e, v);
    return (it != e && *it == v) ? weights[u64(it - col_idx.data())] : INFINITY;
}

CsrGraph CsrGraph::from_rows(std::vector<std::vector<std::pair<u32, f32>>>& rows) {
    CsrGraph g;
    g.n = u32(rows.size());
    g.row_ptr.assign(u64(g.n) + 1, 0);
    for (u32 v = 0; v < g.n; ++v) g.row_ptr[v + 1] = g.row_ptr[v] + rows[v].size();
    const u64 nnz = g.row_ptr[g.n];
    g.m = nnz / 2;
    g.col_idx.resize(nnz);
    g.weights.resize(nnz);
    for (u32 v = 0; v < g.n; ++v) {
        auto& r = rows[v];
        std::sort(r.begin(), r.end());
        u64 o = g.row_ptr[v];
        for (const auto& e : r) {
            g.col_idx[o] = e.first;
            g.weights[o] = e.second;
            ++o;
        }
    }
    return g;
}

bool CsrGraph::valid() const {
    if (row_ptr.size() != u64(n) + 1 || row_ptr[0] != 0 || row_ptr[n] != 2 * m) return false;
    for (u32 v = 0; v < n; ++v)
        for (u64 t = begin(v); t < end(v); ++t) {
            const u32 u = col_idx[t];
            if (u >= n || u == v) return false;
            if (t > begin(v) && col_idx[t - 1] >= u) return false;
            if (weight(u, v) != weights[t]) return false;
        }
    return true;
}

// ---- front-end functions ------------------------------------------------------
static CsrGraph graph_from_abi(u32 n, std::vector<u64>& rp, u64 nnz, u32* col, f32* w) {
    CsrGraph g;
    g.n = n;
    g.row_ptr.swap(rp);
    g.col_idx = take(col, nnz);
    g.weights = take(w, nnz);
    g.m = nnz / 2;
    return g;
}

CsrGraph radius_graph(const PointCloud& c, f32 tau) {
    std::vector<u64> rp(u64(c.n) + 1, 0);
    u64 nnz = 0;
    u32* col = nullptr;
    f32* w = nullptr;
    check(tt_radius_graph(c.x, c.n, c.dim, c.metric.kind, c.metric.p, tau, rp.data(), &nnz, &col, &w));
    return graph_from_abi(c.n, rp, nnz, col, w);
}

KnnLists knn_lists_exact(const PointCloud& c, u32 k) {
    KnnLists L;
    L.n = c.n;
    L.k = c.n ? std::min<u32>(k, c.n - 1) : 0;
    L.ids.resize(u64(L.n) * L.k);
    u32 kk = 0;
    check(tt_knn_exact(c.x, c.n, c.dim, c.metric.kind, c.metric.p, k, &kk, L.ids.data()));
    return L;
}

CsrGraph knn_graph(const PointCloud& c, const KnnLists& L) {
    std::vector<u64> rp(u64(c.n) + 1, 0);
    u64 nnz = 0;
    u32* col = nullptr;
    f32* w = nullptr;
    check(tt_knn_graph(c.x, c.n, c.dim, c.metric.kind, c.metric.p, L.k, L.ids.data(), rp.data(), &nnz,
                       &col, &w));
    return graph_from_abi(c.n, rp, nnz, col, w);
}

namespace mst {
// mst::h0 (mst.hpp:25-99): bars from the sorted Boruvka deaths.
Barcode h0(const PointCloud& c, f32 threshold) {
    Barcode out;
    out.ensure(0);
    std::vector<f32> deaths(c.n > 1 ? c.n - 1 : 0);
    check(tt_mst_deaths(c.x, c.n, c.dim, c.metric.kind, c.metric.p, deaths.data()));
    if (c.n == 0) return out;
    u64 essential = 1;
    for (f32 d : deaths) {
        if (d > threshold) ++essential;
        else if (d > 0.0f) out.add(0, 0.0f, d, 0, 0);
    }
    for (u64 e = 0; e < essential; ++e) out.add(0, 0.0f, INFINITY, 0, cob::kNone<cob::u128>);
    return out;
}
}  // namespace mst

namespace rips {
DenseComplex make_dense(const PointCloud& c, u64 budget) {
    int layout = 0;
    check(tt_dense_layout(c.n, budget, &layout));
    DenseComplex cx;
    cx.n = c.n;
    cx.full = layout == TT_LAYOUT_FULL;
    cx.mat.resize(dense_bytes(c.n, cx.full) / sizeof(f32));
    check(tt_make_dense(c.x, c.n, c.dim, c.metric.kind, c.metric.p, layout, cx.mat.data(), &cx.radius));
    return cx;
}
}  // namespace rips

// ---- persistence ----------------------------------------------------------------
Result persistence(const CsrGraph& g, u32 maxdim, f32 threshold, u32 threads, Mode mode, bool force_wide) {
    const auto t0 = std::chrono::steady_clock::now();
    validate_options(threshold, 1);
    if (!g.valid())
        throw std::invalid_argument("persistence: CsrGraph invariant violated (see CsrGraph::valid)");
    if (threads) set_num_threads(threads);
    Result r;
    r.stats.mode_reason = select_mode(mode, g.n, maxdim, g.n ? 2.0 * double(g.m) / g.n : 0.0);
    r.stats.engine = "coboundary";
    cob::SparseView cx;
    cx.n = g.n;
    cx.threshold = threshold;
    cx.row_ptr = g.row_ptr.data();
    cx.col = g.col_idx.data();
    cx.w = g.weights.data();
    const SparseEdgeList e(g, threshold, maxdim >= 1);
    reduce_all(cx, maxdim, e.idx, e.d, e.apparent, e.m, r.barcode, r.stats, force_wide);
    r.barcode.ensure(maxdim);
    r.stats.method = "sparse";
    r.stats.threshold = threshold;
    r.stats.seconds_total = since(t0);
    r.stats.peak_rss_mb = peak_rss_mb();
    return r;
}

Result persistence(const f32* points, u32 n, u32 dim, const Options& o) {
    const auto t0 = std::chrono::steady_clock::now();
    validate_options(o.threshold, o.graph == Options::Graph::Knn ? o.k : 1);
    if (o.threads) set_num_threads(o.threads);
    const u64 budget = o.memory_budget ? o.memory_budget : default_memory_budget();
    const PointCloud cloud{points, n, dim, Metric::parse(o.metric)};
    for (u64 i = 0; i < u64(n) * dim; ++i)
        if (!std::isfinite(points[i])) throw std::invalid_argument("persistence: non-finite coordinate");
    Result r;
    Stats& st = r.stats;

    if (o.graph == Options::Graph::Full && o.maxdim == 0) {
        r.barcode = mst::h0(cloud, o.threshold);
        st.method = "mst";
        st.engine = "mst";
        st.mode_reason = "H0 only: minimum spanning tree (no reduction engine)";
        st.threshold = o.threshold;
        st.edges = n ? n - 1 : 0;
    } else if (o.graph == Options::Graph::Full && std::isinf(o.threshold)) {
        st.mode_reason = select_mode(o.mode, n, o.maxdim, INFINITY);
        st.engine = "coboundary";
        if (o.maxdim + 2 > 32)
            throw std::invalid_argument(o.mode == Mode::Boundary ? "redzed: maxdim must be <= 30"
                                                                 : "rips: maxdim must be <= 30");
        int layout = 0;
        check(tt_dense_layout(n, budget, &layout));
        if (layout == TT_LAYOUT_FULL && boundary_engine(o.mode, n, o.maxdim, true)) {
            // the matrix is written whole by the library: no zero fill; the
            // edge arrays are used where the library put them
            std::unique_ptr<f32[]> mat(new f32[std::max<u64>(1, u64(n) * n)]);
            u64 m = 0;
            int32_t* pv = nullptr;
            f32* pd = nullptr;
            f32 thr = INFINITY;
            const auto tf = std::chrono::steady_clock::now();
            const tt_status rc = tt_build_edges_witness(points, n, dim, cloud.metric.kind, cloud.metric.p, INFINITY,
                                                        TT_LAYOUT_FULL, mat.get(), &m, &pv, &pd, &thr);
            const std::unique_ptr<int32_t, void (*)(void*)> verts(pv, tt_free);
            const std::unique_ptr<f32, void (*)(void*)> d(pd, tt_free);
            check(rc);
            st.seconds_front = since(tf);
            reduce_boundary(n, mat.get(), m, reinterpret_cast<const u32*>(verts.get()), d.get(), true, r.barcode,
                            st);
            st.engine = "boundary";
            st.method = "dense-redzed";
            st.threshold = thr;
            r.barcode.ensure(o.maxdim);
            st.seconds_total = since(t0);
            st.peak_rss_mb = peak_rss_mb();
            return r;
        }
        std::vector<f32> mat(rips::dense_bytes(n, layout == TT_LAYOUT_FULL) / sizeof(f32));
        u64 m = 0;
        u64* pidx = nullptr;
        f32* pd = nullptr;
        u8* pap = nullptr;
        f32 thr = INFINITY;
        const auto tf = std::chrono::steady_clock::now();
        check(tt_dense_filtration(points, n, dim, cloud.metric.kind, cloud.metric.p, INFINITY, layout,
                                  mat.data(), &m, &pidx, &pd, o.maxdim >= 1 ? &pap : nullptr, &thr));
        st.seconds_front = since(tf);
        std::vector<u64> idx = take(pidx, m);
        std::vector<f32> d = take(pd, m);
        std::vector<u8> ap = take(pap, m);
        cob::DenseView cx;
        cx.n = n;
        cx.threshold = thr;
        cx.full = layout == TT_LAYOUT_FULL;
        cx.mat = mat.data();
        reduce_all(cx, o.maxdim, idx.data(), d.data(), ap.empty() ? nullptr : ap.data(), m, r.barcode, st,
                   o.force_wide);
        st.method = "dense";
        st.threshold = thr;
    } else {
        if (o.approximate) throw std::invalid_argument("persistence: built without HNSW support");
        const auto tg = std::chrono::steady_clock::now();
        CsrGraph g;
        std::vector<u64> rp(u64(n) + 1, 0);
        u64 nnz = 0;
        u32* col = nullptr;
        f32* w = nullptr;
        if (o.graph == Options::Graph::Full)
            check(tt_radius_graph(points, n, dim, cloud.metric.kind, cloud.metric.p, o.threshold, rp.data(),
                                  &nnz, &col, &w));
        else
            check(tt_knn_graph_exact(points, n, dim, cloud.metric.kind, cloud.metric.p, o.k, rp.data(), &nnz,
                                     &col, &w));
        g = graph_from_abi(n, rp, nnz, col, w);
        st.seconds_graph = since(tg);
        const Result rg = persistence(g, o.maxdim, o.threshold, 0, o.mode, o.force_wide);
        r.barcode = rg.barcode;
        const double sg = st.seconds_graph;
        st = rg.stats;
        st.seconds_graph = sg;
        st.method = "sparse";
        st.threshold = o.threshold;
    }
    r.barcode.ensure(o.maxdim);
    st.seconds_total = since(t0);
    st.peak_rss_mb = peak_rss_mb();
    return r;
}

// ---- batches ---------------------------------------------------------------
BatchRunner::BatchRunner(const std::vector<u32>& n, const std::vector<u32>& dim, const std::string& metric,
                         u32 maxdim, Mode mode, bool witness_flags)
    : n_(n), dim_(dim), maxdim_(maxdim) {
    if (n.size() != dim.size()) throw std::invalid_argument("persistence_batch: n and dim differ in length");
    if (maxdim + 2 > 32)
        throw std::invalid_argument(mode == Mode::Boundary ? "redzed: maxdim must be <= 30"
                                                           : "rips: maxdim must be <= 30");
    const Metric m = Metric::parse(metric);
    u32 nmax = 0;
    for (u32 v : n) nmax = std::max(nmax, v);
    (void)select_mode(mode, nmax, maxdim, INFINITY);   // the single-cloud checks and errors
    // maxdim 1: the boundary engine (auto, boundary) or the cohomology engine
    // on the non-apparent edges; maxdim 0: the union-find on the non-apparent
    // edges (an apparent edge never joins two components); maxdim >= 2: the
    // cohomology engine on every kept edge with its apparent flag, since the
    // next dimension is assembled from all of them.
    boundary_ = boundary_engine(mode, nmax, maxdim, true);
    tagged_ = !boundary_ && maxdim >= 2;
    check(tt_batch_create(u32(n.size()), n.data(), dim.data(), m.kind, m.p, INFINITY, &h_));
    if (boundary_) check(tt_batch_set_order(h_, witness_flags ? 2 : 1));
    if (tagged_) check(tt_batch_set_order(h_, 3));
}

BatchRunner::~BatchRunner() {
    if (h_) tt_batch_destroy(h_);
}

u64 BatchRunner::total_points() const {
    u64 t = 0;
    for (size_t c = 0; c < n_.size(); ++c) t += u64(n_[c]) * dim_[c];
    return t;
}

namespace {

// vertices of simplex `idx` with nv vertices, descending (combinatorial
// number system, as Binomials indexes them)
void simplex_vertices(const cob::Binomials<cob::u128>& B, u32 n, cob::u128 idx, u32 nv, u32* out) {
    u32 top = n;
    for (u32 i = 0; i < nv; ++i) {
        const u32 k = nv - i;
        u32 lo = k - 1, hi = top;
        while (hi - lo > 1) {
            const u32 mid = lo + (hi - lo) / 2;
            if (B(mid, k) <= idx) lo = mid;
            else hi = mid;
        }
        out[i] = lo;
        idx -= B(lo, k);
        top = lo;
    }
}

void longest_edge(const cob::Binomials<cob::u128>& B, u32 n, cob::u128 idx, u32 nv,
                  const std::function<f32(u32, u32)>& dist, int64_t* e) {
    u32 vd[34];
    simplex_vertices(B, n, idx, nv, vd);
    f32 best = -1.0f;
    for (u32 a = 0; a < nv; ++a)
        for (u32 b = a + 1; b < nv; ++b) {
            const f32 dd = dist(vd[a], vd[b]);
            if (dd > best) {
                best = dd;
                e[0] = vd[a];
                e[1] = vd[b];
            }
        }
}

}  // namespace

CriticalEdges critical_edges(const Barcode& bc, u32 n, const std::function<f32(u32, u32)>& dist) {
    CriticalEdges ce;
    const u32 dims = u32(bc.birth.size());
    ce.birth.resize(dims);
    ce.death.resize(dims);
    if (!dims) return ce;
    cob::Binomials<cob::u128> B(std::max<u32>(n, 1), dims + 1);
    for (u32 d = 0; d < dims; ++d) {
        const size_t m = bc.birth[d].size();
        ce.birth[d].assign(2 * m, -1);
        ce.death[d].assign(2 * m, -1);
        for (size_t i = 0; i < m; ++i) {
            if (d > 0) longest_edge(B, n, bc.birth_simplex[d][i], d + 1, dist, &ce.birth[d][2 * i]);
            if (std::isfinite(bc.death[d][i]))
                longest_edge(B, n, bc.death_simplex[d][i], d + 2, dist, &ce.death[d][2 * i]);
        }
    }
    return ce;
}

float BatchRunner::time_distance_stage() {
    float ms = 0;
    check(tt_batch_time_dist(h_, &ms));
    return ms;
}

void BatchRunner::front(const f32* pts, bool use_graph) {
    const u64 total = total_points();
    for (u64 i = 0; i < total; ++i)
        if (!std::isfinite(pts[i])) throw std::invalid_argument("persistence: non-finite coordinate");
    check(tt_batch_run(h_, pts, use_graph ? 1 : 0));
    for (u32 c = 0; c < u32(n_.size()); ++c) check(tt_batch_wait(h_, c));
    check(tt_batch_info(h_, &device_ms_, &rpw_, &edges_total_, &edges_kept_));
}

std::vector<Result> BatchRunner::run(const f32* pts, bool use_graph, u32 threads,
                                     std::vector<CriticalEdges>* edges) {
    const auto t0 = std::chrono::steady_clock::now();
    const u64 total = total_points();
    for (u64 i = 0; i < total; ++i)
        if (!std::isfinite(pts[i])) throw std::invalid_argument("persistence: non-finite coordinate");
    check(tt_batch_run(h_, pts, use_graph ? 1 : 0));
    check(tt_batch_info(h_, &device_ms_, &rpw_, &edges_total_, &edges_kept_));
    const u32 nc = u32(n_.size());
    std::vector<Result> out(nc);
    if (edges) edges->assign(nc, {});
    std::vector<u32> order(nc);
    for (u32 c = 0; c < nc; ++c) order[c] = c;
    std::stable_sort(order.begin(), order.end(), [&](u32 a, u32 b) { return n_[a] > n_[b]; });
    std::atomic<u32> next{0};
    std::vector<std::string> errors(nc);
    auto worker = [&] {
        for (u32 k; (k = next.fetch_add(1)) < nc;) {
            const u32 c = order[k];
            try {
                check(tt_batch_wait(h_, c));
                const f32* mat = nullptr;
                f32 thr = 0;
                u64 m = 0;
                const f32* d = nullptr;
                Result& r = out[c];
                if (boundary_) {
                    const u32* verts = nullptr;
                    int witness = 0;
                    check(tt_batch_cloud_edges(h_, c, &mat, &thr, &m, &verts, &d, &witness));
                    reduce_boundary(n_[c], mat, m, verts, d, witness != 0, r.barcode, r.stats);
                    r.stats.engine = "boundary";
                } else {
                    const u64* idx = nullptr;
                    const u8* ap = nullptr;
                    check(tt_batch_cloud(h_, c, &mat, &thr, &m, &idx, &d));
                    if (tagged_) check(tt_batch_cloud_apparent(h_, c, &ap));
                    cob::DenseView cx;
                    cx.n = n_[c];
                    cx.threshold = thr;
                    cx.full = true;
                    cx.mat = mat;
                    reduce_all(cx, maxdim_, idx, d, ap, m, r.barcode, r.stats);
                    r.stats.engine = "coboundary";
                }
                r.barcode.ensure(maxdim_);
                r.stats.method = "dense-batch";
                r.stats.threshold = thr;
                if (edges) {
                    const u64 nn = n_[c];
                    (*edges)[c] = critical_edges(r.barcode, n_[c],
                                                 [&](u32 a, u32 b) { return mat[u64(a) * nn + b]; });
                }
            } catch (const std::exception& e) {
                errors[c] = e.what();
            }
        }
    };
    const u32 nt = std::max<u32>(1, std::min<u32>(threads ? threads : num_threads(), nc));
    std::vector<std::thread> pool;
    for (u32 t = 1; t < nt; ++t) pool.emplace_back(worker);
    worker();
    for (auto& th : pool) th.join();
    for (u32 c = 0; c < nc; ++c)
        if (!errors[c].empty()) throw std::runtime_error("cloud " + std::to_string(c) + ": " + errors[c]);
    const double secs = since(t0);
    for (auto& r : out) r.stats.seconds_total = secs;
    return out;
}

// ---- batched kNN graphs --------------------------------------------------------
BatchKnnRunner::BatchKnnRunner(const std::vector<u32>& n, const std::vector<u32>& dim, const std::string& metric,
                               u32 k, u32 maxdim, Mode mode)
    : n_(n), dim_(dim), maxdim_(maxdim), mode_(mode) {
    if (n.size() != dim.size()) throw std::invalid_argument("persistence_batch: n and dim differ in length");
    const Metric m = Metric::parse(metric);
    check(tt_batch_knn_create(u32(n.size()), n.data(), dim.data(), m.kind, m.p, k, &h_));
}

BatchKnnRunner::~BatchKnnRunner() {
    if (h_) tt_batch_knn_destroy(h_);
}

u64 BatchKnnRunner::total_points() const {
    u64 t = 0;
    for (size_t c = 0; c < n_.size(); ++c) t += u64(n_[c]) * dim_[c];
    return t;
}

void BatchKnnRunner::front(const f32* pts, bool use_graph) {
    check(tt_batch_knn_run(h_, pts, use_graph ? 1 : 0));
    check(tt_batch_knn_info(h_, &device_ms_, &nnz_));
}

CsrGraph BatchKnnRunner::graph(u32 cloud) const {
    if (cloud >= n_.size()) throw std::invalid_argument("persistence_batch: bad cloud");
    const u32* ro = nullptr;
    const u32* col = nullptr;
    const f32* w = nullptr;
    check(tt_batch_knn_cloud(h_, cloud, &ro, &col, &w));
    CsrGraph g;
    g.n = n_[cloud];
    g.row_ptr.resize(u64(g.n) + 1);
    for (u32 v = 0; v <= g.n; ++v) g.row_ptr[v] = u64(ro[v] - ro[0]);
    g.col_idx.assign(col + ro[0], col + ro[g.n]);
    g.weights.assign(w + ro[0], w + ro[g.n]);
    g.m = g.col_idx.size() / 2;
    return g;
}

KnnLists BatchKnnRunner::lists(u32 cloud) {
    if (cloud >= n_.size()) throw std::invalid_argument("persistence_batch: bad cloud");
    KnnLists L;
    const u32* ids = nullptr;
    check(tt_batch_knn_lists(h_, cloud, &L.k, &ids));
    L.n = n_[cloud];
    L.ids.assign(ids, ids + u64(L.n) * L.k);
    return L;
}

std::vector<Result> BatchKnnRunner::run(const f32* pts, bool use_graph, u32 threads,
                                        std::vector<CriticalEdges>* edges) {
    const auto t0 = std::chrono::steady_clock::now();
    front(pts, use_graph);
    const double sg = since(t0);
    const u32 nc = u32(n_.size());
    std::vector<Result> out(nc);
    if (edges) edges->assign(nc, {});
    std::vector<u32> order(nc);
    for (u32 c = 0; c < nc; ++c) order[c] = c;
    std::stable_sort(order.begin(), order.end(), [&](u32 a, u32 b) { return n_[a] > n_[b]; });
    std::atomic<u32> next{0};
    std::vector<std::string> errors(nc);
    auto worker = [&] {
        for (u32 k; (k = next.fetch_add(1)) < nc;) {
            const u32 c = order[k];
            try {
                const CsrGraph g = graph(c);
                Result& r = out[c];
                r = persistence(g, maxdim_, INFINITY, 0, mode_);
                r.barcode.ensure(maxdim_);
                r.stats.method = "sparse-batch";
                r.stats.seconds_graph = sg;
                if (edges) (*edges)[c] = critical_edges(r.barcode, n_[c], [&](u32 a, u32 b) { return g.weight(a, b); });
            } catch (const std::exception& e) {
                errors[c] = e.what();
            }
        }
    };
    const u32 nt = std::max<u32>(1, std::min<u32>(threads ? threads : num_threads(), nc));
    std::vector<std::thread> pool;
    for (u32 t = 1; t < nt; ++t) pool.emplace_back(worker);
    worker();
    for (auto& th : pool) th.join();
    for (u32 c = 0; c < nc; ++c)
        if (!errors[c].empty()) throw std::runtime_error("cloud " + std::to_string(c) + ": " + errors[c]);
    const double secs = since(t0);
    for (auto& r : out) r.stats.seconds_total = secs;
    return out;
}

}  // namespace tops
