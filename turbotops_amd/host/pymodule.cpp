// SPDX-License-Identifier: MIT
// pymodule.cpp — Python module _turbotops, the counterpart of the reference's
// nanobind extension (python/turbotops_ext.cpp:20-120): persistence_points
// and persistence_csr with the same arguments, return values and errors,
// plus persistence_pairs, which also returns each bar's critical edges.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <algorithm>

#include "turbotops_gpu.hpp"

namespace py = pybind11;
using namespace tops;

namespace {

using F32Array = py::array_t<float, py::array::c_style | py::array::forcecast>;
using I64Array = py::array_t<int64_t, py::array::c_style | py::array::forcecast>;

// bars of each dimension as an (m, 2) float32 array (turbotops_ext.cpp:20-33)
py::list to_python(const Barcode& bc) {
    py::list out;
    for (size_t d = 0; d < bc.birth.size(); ++d) {
        const size_t m = bc.birth[d].size();
        py::array_t<float> a({m, size_t(2)});
        auto v = a.mutable_unchecked<2>();
        for (size_t i = 0; i < m; ++i) {
            v(i, 0) = bc.birth[d][i];
            v(i, 1) = bc.death[d][i];
        }
        out.append(a);
    }
    return out;
}

// stats_dict (turbotops_ext.cpp:35-51)
py::dict stats_dict(const Stats& s) {
    py::dict d;
    d["method"] = s.method;
    d["threshold"] = s.threshold;
    d["edges"] = s.edges;
    d["index_bits"] = s.wide_index ? 128 : 64;
    d["seconds_graph"] = s.seconds_graph;
    d["seconds_total"] = s.seconds_total;
    d["peak_rss_mb"] = s.peak_rss_mb;
    d["seconds_front"] = s.seconds_front;
    d["seconds_reduce"] = s.seconds_reduce;
    d["engine"] = s.engine;
    return d;
}

Options make_options(u32 maxdim, float threshold, const std::string& metric, const std::string& graph, u32 k,
                     bool approximate, u64 recall_samples, u32 threads, u64 memory_budget,
                     const std::string& mode = "auto", bool force_wide = false) {
    Options o;
    o.mode = parse_mode(mode);
    o.force_wide = force_wide;
    o.maxdim = maxdim;
    o.threshold = threshold;
    o.metric = metric;
    if (graph == "full") o.graph = Options::Graph::Full;
    else if (graph == "knn") o.graph = Options::Graph::Knn;
    else throw py::value_error("graph must be 'full' or 'knn'");
    o.k = k;
    o.approximate = approximate;
    o.recall_samples = recall_samples;
    o.threads = threads;
    o.memory_budget = memory_budget;
    return o;
}

Result run_points(F32Array x, const Options& o) {
    if (x.ndim() != 2) throw py::value_error("x must be a 2-d array (n, d)");
    const u32 n = u32(x.shape(0)), dim = u32(x.shape(1));
    Result r;
    {
        py::gil_scoped_release release;
        r = persistence(x.data(), n, dim, o);
    }
    return r;
}

// mode and force_wide are not in the reference's Python API: the
// engine choice (--mode of its CLI) and rips::compute's force_wide, for tests.
py::tuple persistence_points(F32Array x, u32 maxdim, float threshold, const std::string& metric,
                             const std::string& graph, u32 k, bool approximate, u64 recall_samples, u32 threads,
                             u64 memory_budget, const std::string& mode, bool force_wide) {
    const Result r = run_points(x, make_options(maxdim, threshold, metric, graph, k, approximate, recall_samples,
                                                threads, memory_budget, mode, force_wide));
    return py::make_tuple(to_python(r.barcode), stats_dict(r.stats));
}

// persistence_csr (turbotops_ext.cpp:76-106): both directions of every
// entry, the smallest weight of repeated entries, then persistence(CsrGraph).
py::tuple persistence_csr(I64Array indptr, I64Array indices, F32Array data, u32 n, u32 maxdim, float threshold,
                          u32 threads, bool force_wide) {
    if (size_t(indptr.shape(0)) != size_t(n) + 1 || indices.shape(0) != data.shape(0))
        throw py::value_error("inconsistent CSR arrays");
    std::vector<std::vector<std::pair<u32, f32>>> rows(n);
    const int64_t* ip = indptr.data();
    const int64_t* ix = indices.data();
    const float* dv = data.data();
    for (u32 i = 0; i < n; ++i)
        for (int64_t t = ip[i]; t < ip[i + 1]; ++t) {
            const int64_t j = ix[t];
            const float w = dv[t];
            if (j < 0 || j >= int64_t(n)) throw py::value_error("column index out of range");
            if (!(w >= 0) || !std::isfinite(w)) throw py::value_error("weights must be finite and >= 0");
            if (j == i) continue;
            rows[i].emplace_back(u32(j), w);
            rows[u32(j)].emplace_back(i, w);
        }
    for (auto& r : rows) {
        std::sort(r.begin(), r.end());
        r.erase(std::unique(r.begin(), r.end(), [](const auto& a, const auto& b) { return a.first == b.first; }),
                r.end());
    }
    const CsrGraph g = CsrGraph::from_rows(rows);
    Result r;
    {
        py::gil_scoped_release release;
        r = persistence(g, maxdim, threshold, threads, Mode::Auto, force_wide);
    }
    return py::make_tuple(to_python(r.barcode), stats_dict(r.stats));
}

// critical edges of one dimension as (m, 2) int64 arrays
py::array_t<int64_t> edge_array(const std::vector<int64_t>& flat) {
    const size_t m = flat.size() / 2;
    py::array_t<int64_t> a({m, size_t(2)});
    std::copy(flat.begin(), flat.end(), a.mutable_data());
    return a;
}

py::tuple edges_to_python(const CriticalEdges& ce) {
    py::list births, deaths;
    for (size_t d = 0; d < ce.birth.size(); ++d) {
        births.append(edge_array(ce.birth[d]));
        deaths.append(edge_array(ce.death[d]));
    }
    return py::make_tuple(births, deaths);
}

// persistence on a point cloud, returning per dimension the bars and, per
// bar, the critical edges of its birth and death simplices ((-1, -1) where
// there is none: H0 births, essential deaths).
py::tuple persistence_pairs(F32Array x, u32 maxdim, float threshold, const std::string& metric,
                            const std::string& graph, u32 k, u32 threads, u64 memory_budget) {
    const Options o = make_options(maxdim, threshold, metric, graph, k, false, 0, threads, memory_budget);
    if (o.graph == Options::Graph::Full && maxdim == 0)
        throw py::value_error("persistence_pairs: maxdim 0 runs as an MST; use maxdim >= 1");
    const Result r = run_points(x, o);
    const u32 n = u32(x.shape(0)), dim = u32(x.shape(1));
    const Metric mt = Metric::parse(metric);
    // the filtration's own distances, for the tie-break between equally long edges
    std::vector<float> full(size_t(n) * n);
    float radius = 0;
    check(tt_make_dense(x.data(), n, dim, mt.kind, mt.p, TT_LAYOUT_FULL, full.data(), &radius));
    const CriticalEdges ce = critical_edges(r.barcode, n, [&](u32 a, u32 b) { return full[size_t(a) * n + b]; });
    const py::tuple e = edges_to_python(ce);
    return py::make_tuple(to_python(r.barcode), e[0], e[1], stats_dict(r.stats));
}

// Batch runners: points as one packed float32 array (clouds concatenated)
// or a list of (n_i, d_i) arrays.
template <class R>
std::vector<float> pack_points(const R& br, py::object clouds) {
    std::vector<float> packed;
    if (py::isinstance<py::list>(clouds) || py::isinstance<py::tuple>(clouds)) {
        for (auto h : clouds) {
            F32Array a = py::cast<F32Array>(h);
            packed.insert(packed.end(), a.data(), a.data() + a.size());
        }
    } else {
        F32Array a = py::cast<F32Array>(clouds);
        packed.assign(a.data(), a.data() + a.size());
    }
    if (packed.size() != br.total_points()) throw py::value_error("clouds do not match the batch shapes");
    return packed;
}

template <class R>
void batch_front(R& br, py::object clouds, bool use_graph) {
    const std::vector<float> packed = pack_points(br, clouds);
    py::gil_scoped_release release;
    br.front(packed.data(), use_graph);
}

py::tuple knn_batch_graph(const BatchKnnRunner& br, u32 cloud) {
    const CsrGraph g = br.graph(cloud);
    py::array_t<uint64_t> rp(g.row_ptr.size());
    std::copy(g.row_ptr.begin(), g.row_ptr.end(), rp.mutable_data());
    py::array_t<uint32_t> col(g.col_idx.size());
    std::copy(g.col_idx.begin(), g.col_idx.end(), col.mutable_data());
    py::array_t<float> w(g.weights.size());
    std::copy(g.weights.begin(), g.weights.end(), w.mutable_data());
    return py::make_tuple(rp, col, w);
}

py::array_t<uint32_t> knn_batch_lists(BatchKnnRunner& br, u32 cloud) {
    const KnnLists L = br.lists(cloud);
    py::array_t<uint32_t> a({py::ssize_t(L.n), py::ssize_t(L.k)});
    std::copy(L.ids.begin(), L.ids.end(), a.mutable_data());
    return a;
}


// With pairs, each entry also carries the cloud's critical edges:
// (dgms, birth_edges, death_edges, info) as persistence_pairs.
template <class R>
py::list batch_run(R& br, py::object clouds, bool use_graph, u32 threads, bool pairs) {
    const std::vector<float> packed = pack_points(br, clouds);
    const float* p = packed.data();
    std::vector<Result> rs;
    std::vector<CriticalEdges> ce;
    {
        py::gil_scoped_release release;
        rs = br.run(p, use_graph, threads, pairs ? &ce : nullptr);
    }
    py::list out;
    for (size_t c = 0; c < rs.size(); ++c) {
        if (pairs) {
            const py::tuple e = edges_to_python(ce[c]);
            out.append(py::make_tuple(to_python(rs[c].barcode), e[0], e[1], stats_dict(rs[c].stats)));
        } else {
            out.append(py::make_tuple(to_python(rs[c].barcode), stats_dict(rs[c].stats)));
        }
    }
    return out;
}

}  // namespace

PYBIND11_MODULE(_turbotops, m) {
    m.doc() = "Vietoris-Rips persistent homology with an MI355X front end";
    m.attr("__version__") = "0.1.0";
    m.def("persistence_points", &persistence_points, py::arg("x"), py::arg("maxdim"), py::arg("threshold"),
          py::arg("metric"), py::arg("graph"), py::arg("k"), py::arg("approximate"), py::arg("recall_samples"),
          py::arg("threads"), py::arg("memory_budget"), py::arg("mode") = "auto", py::arg("force_wide") = false);
    m.def("persistence_csr", &persistence_csr, py::arg("indptr"), py::arg("indices"), py::arg("data"),
          py::arg("n"), py::arg("maxdim"), py::arg("threshold"), py::arg("threads"), py::arg("force_wide") = false);
    m.def("persistence_pairs", &persistence_pairs, py::arg("x"), py::arg("maxdim"), py::arg("threshold"),
          py::arg("metric"), py::arg("graph"), py::arg("k"), py::arg("threads"), py::arg("memory_budget"));
    py::class_<BatchRunner>(m, "BatchRunner")
        .def(py::init([](const std::vector<u32>& n, const std::vector<u32>& dim, const std::string& metric,
                         u32 maxdim, const std::string& mode, bool witness_flags) {
                 return std::make_unique<BatchRunner>(n, dim, metric, maxdim, parse_mode(mode), witness_flags);
             }),
             py::arg("n"), py::arg("dim"), py::arg("metric") = "l2", py::arg("maxdim") = 1,
             py::arg("mode") = "auto", py::arg("witness_flags") = true)
        .def("front", &batch_front<BatchRunner>, py::arg("clouds"), py::arg("use_graph") = true)
        .def("run", &batch_run<BatchRunner>, py::arg("clouds"), py::arg("use_graph") = true, py::arg("threads") = 0,
             py::arg("pairs") = false)
        .def("time_distance_stage", &BatchRunner::time_distance_stage)
        .def_property_readonly("device_ms", &BatchRunner::device_ms)
        .def_property_readonly("rows_per_wave", &BatchRunner::rows_per_wave)
        .def_property_readonly("edges_total", &BatchRunner::edges_total)
        .def_property_readonly("edges_kept", &BatchRunner::edges_kept);
    py::class_<BatchKnnRunner>(m, "BatchKnnRunner")
        .def(py::init([](const std::vector<u32>& n, const std::vector<u32>& dim, const std::string& metric, u32 k,
                         u32 maxdim, const std::string& mode) {
                 return std::make_unique<BatchKnnRunner>(n, dim, metric, k, maxdim, parse_mode(mode));
             }),
             py::arg("n"), py::arg("dim"), py::arg("metric") = "l2", py::arg("k") = 15, py::arg("maxdim") = 1,
             py::arg("mode") = "auto")
        .def("front", &batch_front<BatchKnnRunner>, py::arg("clouds"), py::arg("use_graph") = true)
        .def("run", &batch_run<BatchKnnRunner>, py::arg("clouds"), py::arg("use_graph") = true, py::arg("threads") = 0,
             py::arg("pairs") = false)
        .def("graph", &knn_batch_graph, py::arg("cloud"))
        .def("lists", &knn_batch_lists, py::arg("cloud"))
        .def_property_readonly("device_ms", &BatchKnnRunner::device_ms)
        .def_property_readonly("nnz", &BatchKnnRunner::nnz);
    py::register_exception_translator([](std::exception_ptr p) {
        try {
            if (p) std::rethrow_exception(p);
        } catch (const std::length_error& e) {
            PyErr_SetString(PyExc_ValueError, e.what());
        }
    });
}
