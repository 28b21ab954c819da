// SPDX-License-Identifier: MIT
// turbotops_gpu.hpp — the reference's C++ entry points (dnbaker/turbotops
// v0.1.0, include/turbotops/turbotops.hpp) with every data-parallel front-end
// stage running on an MI355X through the C ABI in include/ttgpu.h.
//
// Same names, argument meaning and exceptions as the reference:
//   tops::persistence(points, n, dim, Options)      turbotops.hpp:211-303
//   tops::persistence(CsrGraph, maxdim, threshold)  turbotops.hpp:179-209
//   tops::radius_graph / knn_lists_exact / knn_graph (graph.hpp:170-304)
//   tops::mst::h0                                    (mst.hpp:25-99)
//   tops::rips::make_dense / DenseComplex            (rips.hpp:255-416)
// Persistence pairs come from the cohomology engine (cohomology.hpp: the
// column reduction on the host, its per-simplex work on the GPU) or, for
// dense H1, the boundary engine (zerodiff.hpp).
#pragma once

#include <cmath>
#include <cstdint>
#include <functional>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "cohomology.hpp"
#include "ttgpu.h"

namespace tops {

using u8 = uint8_t;
using u32 = uint32_t;
using u64 = uint64_t;
using f32 = float;

// Throws the reference's exception class for a C ABI status.
inline void check(tt_status st) {
    if (st == TT_OK) return;
    const std::string m = tt_last_error();
    switch (st) {
        case TT_EINVAL: throw std::invalid_argument(m);
        case TT_ELENGTH: throw std::length_error(m);
        case TT_EOVERFLOW: throw std::overflow_error(m);
        default: throw std::runtime_error(m);
    }
}

// Metric::parse (metric.hpp:214-267).
struct Metric {
    enum Kind { L1 = 0, L2 = 1, Linf = 2, Lp = 3 } kind = L2;
    f32 p = 2.0f;
    static Metric parse(const std::string& s) {
        int k = 0;
        float p = 0;
        check(tt_metric_parse(s.c_str(), &k, &p));
        Metric m;
        m.kind = Kind(k);
        m.p = p;
        return m;
    }
};

// PointCloud (metric.hpp:270-278): borrowed row-major coordinates.
struct PointCloud {
    const f32* x = nullptr;
    u32 n = 0, dim = 0;
    Metric metric;
};

// CsrGraph (core.hpp:92-149).
struct CsrGraph {
    u32 n = 0;
    u64 m = 0;
    std::vector<u64> row_ptr;
    std::vector<u32> col_idx;
    std::vector<f32> weights;

    u64 begin(u32 v) const { return row_ptr[v]; }
    u64 end(u32 v) const { return row_ptr[v + 1]; }
    u64 degree(u32 v) const { return end(v) - begin(v); }
    f32 weight(u32 u, u32 v) const;
    static CsrGraph from_rows(std::vector<std::vector<std::pair<u32, f32>>>& rows);
    bool valid() const;
};

// KnnLists (graph.hpp:205-209).
struct KnnLists {
    u32 n = 0, k = 0;
    std::vector<u32> ids;
    const u32* row(u32 i) const { return ids.data() + u64(i) * k; }
};

using Barcode = cob::Diagram;

CsrGraph radius_graph(const PointCloud& c, f32 tau);
KnnLists knn_lists_exact(const PointCloud& c, u32 k);
CsrGraph knn_graph(const PointCloud& c, const KnnLists& L);

namespace mst {
Barcode h0(const PointCloud& c, f32 threshold = INFINITY);
}

namespace rips {
// DenseComplex (rips.hpp:262-381): the matrix in the full or lower layout.
struct DenseComplex {
    u32 n = 0;
    f32 threshold = INFINITY;
    bool full = false;
    std::vector<f32> mat;
    f32 radius = INFINITY;   // enclosing radius, computed with the matrix
    f32 dist(u32 i, u32 j) const {
        if (i == j) return 0.0f;
        if (full) return mat[u64(i) * n + j];
        if (i < j) std::swap(i, j);
        return mat[u64(i) * (i - 1) / 2 + j];
    }
    f32 enclosing_radius() const { return radius; }
};
inline u64 dense_bytes(u32 n, bool full) {
    return full ? u64(n) * n * sizeof(f32) : u64(n) * (n > 0 ? n - 1 : 0) / 2 * sizeof(f32);
}
DenseComplex make_dense(const PointCloud& c, u64 budget_bytes = ~0ull);
}  // namespace rips

enum class Mode { Auto, Boundary, Coboundary };
Mode parse_mode(const std::string& s);

struct Options {
    u32 maxdim = 1;
    Mode mode = Mode::Auto;
    f32 threshold = INFINITY;
    std::string metric = "l2";
    enum class Graph { Full, Knn } graph = Graph::Full;
    u32 k = 15;
    bool approximate = false;
    u64 recall_samples = 0;
    u64 memory_budget = 0;   // bytes; 0 = half of physical memory
    u32 threads = 0;         // 0 = hardware concurrency
    // 128-bit simplex indices in the cohomology engine even where 64 bits
    // suffice, as the force_wide argument of rips::compute
    // (rips.hpp:1038-1061); for tests
    bool force_wide = false;
};

struct Stats {
    std::string method;
    std::string engine;
    std::string mode_reason;
    f32 threshold = INFINITY;
    u64 edges = 0;
    bool wide_index = false;
    double seconds_graph = 0, seconds_front = 0, seconds_reduce = 0, seconds_total = 0;
    double peak_rss_mb = 0;
    std::vector<u64> columns, apparent, emergent;
};

struct Result {
    Barcode barcode;
    Stats stats;
};

Result persistence(const f32* points, u32 n, u32 dim, const Options& o);
Result persistence(const CsrGraph& g, u32 maxdim, f32 threshold = INFINITY, u32 threads = 0,
                   Mode mode = Mode::Auto, bool force_wide = false);

// Critical edges of a diagram: per dimension and bar, the longest edge
// (a, b) of its birth simplex and of its death simplex, read with dist(a, b)
// (ties: the first pair in descending vertex order); -1 where there is none
// (H0 births, essential deaths). Every finite bar value is the length of
// its critical edge, which is what the backward of a diagram needs.
struct CriticalEdges {
    std::vector<std::vector<int64_t>> birth, death;   // per dimension, 2 per bar
};
CriticalEdges critical_edges(const Barcode& bc, u32 n, const std::function<f32(u32, u32)>& dist);

// Many clouds through one batched front end (include/ttgpu.h tt_batch_*):
// dense full Rips at each cloud's enclosing radius, maxdim 1. Create once per
// set of shapes; run() reduces the clouds on host threads (largest first),
// each as soon as its data has landed. Results equal persistence() on each
// cloud alone.
class BatchRunner {
  public:
    // mode: at maxdim 1, Auto and Boundary reduce with the boundary engine
    // (every kept edge in (diameter, comb index) order), Coboundary with the
    // cohomology engine (the non-apparent edges in its order); maxdim 0 and
    // >= 2 run the cohomology engine (maxdim >= 2 on every kept edge with its
    // apparent flag). witness_flags: the
    // boundary engine's inputs carry the GPU's ready-witness flags (faster
    // host reductions on clouds without long cycles, one more device scan;
    // used where the flag bit fits the batch's edge keys).
    BatchRunner(const std::vector<u32>& n, const std::vector<u32>& dim, const std::string& metric,
                u32 maxdim = 1, Mode mode = Mode::Auto, bool witness_flags = true);
    ~BatchRunner();
    BatchRunner(const BatchRunner&) = delete;
    BatchRunner& operator=(const BatchRunner&) = delete;
    // With `edges` != nullptr, also each cloud's critical edges.
    std::vector<Result> run(const f32* packed_points, bool use_graph = true, u32 threads = 0,
                            std::vector<CriticalEdges>* edges = nullptr);
    // The device front end alone (distances, radii, edge sort, apparent
    // scan) and the copies of its results to host memory, no reduction.
    void front(const f32* packed_points, bool use_graph = true);
    float device_ms() const { return device_ms_; }
    // the distance stage alone (tt_batch_time_dist), in ms
    float time_distance_stage();
    int rows_per_wave() const { return rpw_; }
    u64 edges_total() const { return edges_total_; }
    u64 edges_kept() const { return edges_kept_; }
    u64 total_points() const;

  private:
    tt_batch* h_ = nullptr;
    std::vector<u32> n_, dim_;
    u32 maxdim_ = 1;
    bool boundary_ = true;
    bool tagged_ = false;   // maxdim >= 2: every kept edge with its apparent flag (order 3)
    float device_ms_ = 0;
    int rpw_ = 0;
    u64 edges_total_ = 0, edges_kept_ = 0;
};

// The batched kNN-graph front end (tt_batch_knn_*): every cloud's exact
// kNN lists and symmetric graph in one device sequence, then the sparse
// reductions of the graphs on host threads (largest first). Results equal
// persistence(cloud, graph = knn, k) on each cloud alone.
class BatchKnnRunner {
  public:
    BatchKnnRunner(const std::vector<u32>& n, const std::vector<u32>& dim, const std::string& metric, u32 k,
                   u32 maxdim = 1, Mode mode = Mode::Auto);
    ~BatchKnnRunner();
    BatchKnnRunner(const BatchKnnRunner&) = delete;
    BatchKnnRunner& operator=(const BatchKnnRunner&) = delete;
    std::vector<Result> run(const f32* packed_points, bool use_graph = true, u32 threads = 0,
                            std::vector<CriticalEdges>* edges = nullptr);
    // the device stage alone; the graphs land in host memory
    void front(const f32* packed_points, bool use_graph = true);
    CsrGraph graph(u32 cloud) const;   // after front() or run()
    KnnLists lists(u32 cloud);
    float device_ms() const { return device_ms_; }
    u64 nnz() const { return nnz_; }
    u64 total_points() const;

  private:
    tt_batch_knn* h_ = nullptr;
    std::vector<u32> n_, dim_;
    u32 maxdim_ = 1;
    Mode mode_ = Mode::Auto;
    float device_ms_ = 0;
    u64 nnz_ = 0;
};

void set_num_threads(u32 t);
u32 num_threads();
u64 default_memory_budget();
double peak_rss_mb();

}  // namespace tops
