/* SPDX-License-Identifier: MIT
 * ttgpu.h — C ABI of the MI355X (gfx950) filtration front end.
 *
 * Every entry point replaces one data-parallel function of the reference
 * library (dnbaker/turbotops v0.1.0, header-only C++20); the comment above
 * each declaration cites the reference interface it stands in for.
 *
 * Conventions
 *   - Return value: tt_status. On failure, tt_last_error() returns a
 *     thread-local message whose text matches the reference's exception
 *     message where the reference has one. Status codes map to the
 *     reference's exception classes: TT_EINVAL -> std::invalid_argument,
 *     TT_ELENGTH -> std::length_error, TT_EOVERFLOW -> std::overflow_error,
 *     TT_EDEVICE -> std::runtime_error (HIP failure; never a silent CPU
 *     fallback).
 *   - Metric kinds follow the reference's Metric::Kind
 *     (include/turbotops/metric.hpp:215): 0 L1, 1 L2, 2 Linf, 3 Lp(p).
 *   - Points are row-major n x dim float32.
 *   - "_dev" entry points take device pointers and run asynchronously on
 *     `stream` (a hipStream_t; NULL = the null stream). Host entry points
 *     take host pointers, are synchronous and thread-safe.
 *   - Results do not depend on tile sizes or launch shapes.
 */
#ifndef TTGPU_H
#define TTGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    TT_OK = 0,
    TT_EINVAL = 1,
    TT_ELENGTH = 2,
    TT_EOVERFLOW = 3,
    TT_EDEVICE = 4
} tt_status;

enum { TT_METRIC_L1 = 0, TT_METRIC_L2 = 1, TT_METRIC_LINF = 2, TT_METRIC_LP = 3 };
enum { TT_LAYOUT_FULL = 0, TT_LAYOUT_LOWER = 1 };

/* ---- library ------------------------------------------------------------ */

/* Thread-local message of the last failed call on this thread. */
const char* tt_last_error(void);
/* Library version, "MAJOR.MINOR.PATCH" (the reference's
 * TURBOTOPS_VERSION_STRING it is a drop-in for, core.hpp:8-11). */
const char* tt_version(void);
/* Number of visible HIP devices (0 if none); no device is initialised by
 * loading the library. */
int tt_device_count(void);

/* Metric::parse (metric.hpp:240-253) and Metric::minkowski (:227-237):
 * "l1" | "manhattan" | "l2" | "euclidean" | "linf" | "chebyshev" |
 * "lp:<p>" | "lp:inf"; p < 1 or NaN -> TT_EINVAL "metric: p must be in
 * [1, inf]"; unknown -> TT_EINVAL "metric: unknown '<s>' (...)". */
tt_status tt_metric_parse(const char* metric, int* kind, float* p);

/* ---- dense distance matrix + enclosing radius ---------------------------- */

/* dense_bytes (rips.hpp:384-386) and the layout choice of make_dense
 * (rips.hpp:395, :404, :410-414): full when 4n^2 <= budget, else lower when
 * 2n(n-1) <= budget, else TT_ELENGTH with the reference's message. */
tt_status tt_dense_layout(uint32_t n, uint64_t budget_bytes, int* layout);

/* rips::make_dense(n, PointCloud, budget) (rips.hpp:391-416) plus
 * DenseComplex::enclosing_radius() (rips.hpp:370-380), fused.
 * mat: host buffer of n*n (full) or n(n-1)/2 (lower) floats, written exactly
 * as the reference lays them out. radius may be NULL. */
tt_status tt_make_dense(const float* pts, uint32_t n, uint32_t dim, int kind, float p,
                        int layout, float* mat, float* radius);

/* Device form of tt_make_dense. d_rowmax (n uint32, may be NULL) receives
 * max_j d(i, j) as float bits; d_radius (1 float, may be NULL) receives
 * min_i of those, i.e. the enclosing radius. */
tt_status tt_dist_matrix_dev(const float* d_pts, uint32_t n, uint32_t dim, int kind, float p,
                             int layout, float* d_mat, uint32_t* d_rowmax, float* d_radius,
                             void* stream);

/* tt_dist_matrix_dev with a measurement hook (no counterpart in the
 * reference: for benchmarks that time the distance kernel by itself).
 * ev_begin / ev_end are hipEvent_t handles (either may be NULL), recorded on
 * `stream` immediately before and after the distance-tile launch, i.e.
 * around that one kernel and not the row-maxima fill or the radius
 * reduction. n = 0 with an event gives TT_EINVAL (nothing is launched). */
tt_status tt_dist_matrix_dev_events(const float* d_pts, uint32_t n, uint32_t dim, int kind, float p,
                                    int layout, float* d_mat, uint32_t* d_rowmax, float* d_radius,
                                    void* stream, void* ev_begin, void* ev_end);

/* ---- neighbourhood graphs ----------------------------------------------- */

/* radius_graph(PointCloud, tau) (graph.hpp:170-201) with
 * CsrGraph::from_rows (core.hpp:113-134): every pair with d <= tau, each edge
 * in both rows, rows sorted by neighbour id, no self loops. row_ptr: n + 1
 * entries (caller's); *col and *w (*nnz = 2m entries) are malloc'd by the
 * library, release with tt_free. */
tt_status tt_radius_graph(const float* pts, uint32_t n, uint32_t dim, int kind, float p,
                          float tau, uint64_t* row_ptr, uint64_t* nnz, uint32_t** col, float** w);

/* knn_lists_exact(PointCloud, k) (graph.hpp:211-225): per point the k
 * nearest j != i by (distance, id), ascending; k is clamped to n - 1 and
 * written to *k_out. ids: n * k_out uint32 (caller's). Any k: up to 64 the
 * lists are kept per lane in registers, beyond by a per-query selection;
 * n * k_out < 2^32 (TT_ELENGTH beyond). */
tt_status tt_knn_exact(const float* pts, uint32_t n, uint32_t dim, int kind, float p, uint32_t k,
                       uint32_t* k_out, uint32_t* ids);

/* knn_graph(PointCloud, KnnLists) (graph.hpp:286-304): the symmetric union
 * of the lists (ids: n * k, k already clamped) with recomputed weights, as
 * CSR; output buffers as tt_radius_graph. */
tt_status tt_knn_graph(const float* pts, uint32_t n, uint32_t dim, int kind, float p,
                       uint32_t k, const uint32_t* ids, uint64_t* row_ptr, uint64_t* nnz,
                       uint32_t** col, float** w);

/* knn_graph(cloud, knn_lists_exact(cloud, k)) as persistence() calls them
 * (turbotops.hpp:275-277), with the lists kept in HBM. k == 0 -> TT_EINVAL
 * "persistence: k must be >= 1" (turbotops.hpp:174). */
tt_status tt_knn_graph_exact(const float* pts, uint32_t n, uint32_t dim, int kind, float p,
                             uint32_t k, uint64_t* row_ptr, uint64_t* nnz, uint32_t** col,
                             float** w);

/* ---- H0 ------------------------------------------------------------------ */

/* mst::h0(PointCloud, threshold) (mst.hpp:25-99): the n - 1 sorted MST edge
 * weights (the H0 deaths) by Boruvka; bars follow exactly as mst.hpp:91-97.
 * Non-finite coordinates -> TT_EINVAL "mst: non-finite coordinate". */
tt_status tt_mst_deaths(const float* pts, uint32_t n, uint32_t dim, int kind, float p,
                        float* deaths);

/* Device form; *rounds, *scans and *pairs (each may be NULL) report the
 * Boruvka rounds, the vertex searches (each for the nearest vertex outside
 * the vertex's component, over a spatially pruned candidate set) and the
 * point pairs those searches measured. */
tt_status tt_mst_deaths_dev(const float* d_pts, uint32_t n, uint32_t dim, int kind, float p,
                            float* d_deaths, uint32_t* rounds, uint64_t* scans, uint64_t* pairs,
                            void* stream);

/* tt_mst_deaths_dev without any device allocation. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * Every temporary lives in one static 128 MiB scratch region per device,
 * held until the process ends; n <= 1048576 and about n (56.6 + 4.125 dim)
 * bytes of temporaries (dim <= 17 at n = 2^20; TT_ELENGTH beyond). The stream must
 * belong to the current device. Calls that overlap on one device race on
 * that region: serialize them. Same results as tt_mst_deaths_dev. */
tt_status tt_mst_deaths_dev_noalloc(const float* d_pts, uint32_t n, uint32_t dim, int kind, float p,
                                    float* d_deaths, void* stream);

/* ---- dimension 0 of the coboundary reducer -------------------------------- */

/* Reducer::dim0's kept edges and sort (rips.hpp:685-839) and its parallel
 * apparent test (rips.hpp:844-849 -> zero_apparent_cofacet, :585-665), on the
 * dense full-layout complex: every edge j < i with d(i, j) <= threshold as
 * (idx = i(i-1)/2 + j, d), ordered by `precedes` (rips.hpp:116-118:
 * diameter ascending, index descending), and apparent[e] = 1 when the edge
 * is the zero-pivot facet of its zero-pivot cofacet. threshold = +inf means
 * the enclosing radius (rips.hpp:1073), reported in *threshold_out.
 * The library mallocs *idx, *d and (if apparent != NULL) *apparent; release
 * them with tt_free. Any n whose matrix fits in device memory; the sort keys
 * hold the edge index and the diameter's bits above the smallest kept one
 * (TT_ELENGTH in the rare case that index and diameter range need more than
 * 63 bits together). */
tt_status tt_dense_edges(const float* pts, uint32_t n, uint32_t dim, int kind, float p,
                         float threshold, uint64_t* m, uint64_t** idx, float** d,
                         uint8_t** apparent, float* threshold_out);

/* Device form over a full-layout matrix already in HBM. With d_idx == NULL
 * only *m is computed; otherwise the outputs (capacity entries) are filled. */
tt_status tt_dense_edges_dev(const float* d_mat, uint32_t n, float threshold, int want_apparent,
                             uint64_t capacity, uint64_t* m, uint64_t* d_idx, float* d_d,
                             uint8_t* d_apparent, void* stream);

/* The dense front end of rips::compute_dense and persistence()'s dense branch
 * (rips.hpp:391-416 make_dense, :370-380 enclosing_radius, :1073 threshold,
 * :685-849 dim0 sort and apparent test) in one device pass: mat (host, may be
 * NULL) receives the matrix in `layout`; the edges, apparent flags and
 * threshold are as tt_dense_edges. */
tt_status tt_dense_filtration(const float* pts, uint32_t n, uint32_t dim, int kind, float p,
                              float threshold, int layout, float* mat, uint64_t* m,
                              uint64_t** idx, float** d, uint8_t** apparent,
                              float* threshold_out);

/* Reducer::dim0's kept edges, sort and apparent test on a sparse complex:
 * the flag complex of a weighted graph in CSR form (rows sorted by neighbour
 * id, as CsrGraph::from_rows builds them, core.hpp:113-134), as
 * persistence(const CsrGraph&, ...) (turbotops.hpp:179-209) and the kNN /
 * radius-graph branches of persistence() (turbotops.hpp:255-303) feed the
 * reducer (rips.hpp:685-849, SparseComplex::cofacets :428-471). Every edge
 * v > u with w <= threshold as (idx = v(v-1)/2 + u, w) in `precedes` order
 * (diameter ascending, index descending), and apparent[e] = 1 when the edge
 * is the zero-pivot facet of its zero-pivot cofacet. The library mallocs
 * *idx, *d and (if apparent != NULL) *apparent; release them with tt_free. */
tt_status tt_sparse_edges(uint32_t n, const uint64_t* row_ptr, const uint32_t* col, const float* w,
                          float threshold, uint64_t* m, uint64_t** idx, float** d, uint8_t** apparent);

/* ---- columns of the coboundary reducer above dimension 0 ------------------ */

/* A device copy of the complex the coboundary reducer works on, for its
 * per-simplex work above dimension 0. Simplex indices cross this interface
 * as index_words 64-bit words each (1: 64-bit indices, 2: 128-bit, low word
 * first), the width rips::compute picks (rips.hpp:1051-1061: u64 when every
 * index up to maxdim + 2 vertices fits, else u128; force_wide: u128).
 * Dense: the n x n matrix in `layout` (TT_LAYOUT_FULL or TT_LAYOUT_LOWER),
 * host memory, or with on_device (full layout only) a device pointer the
 * complex borrows until it is destroyed. Sparse: a CSR graph (rows sorted by
 * neighbour id, CsrGraph::from_rows, core.hpp:113-134). threshold: the
 * filtration's cap (the enclosing radius for a dense complex at +inf). Fails
 * with TT_EOVERFLOW when C(n, maxdim + 2) does not fit the index width. */
typedef struct tt_complex tt_complex;
tt_status tt_complex_dense(const float* mat, uint32_t n, int layout, int on_device, float threshold,
                           uint32_t maxdim, int index_words, tt_complex** out);
tt_status tt_complex_sparse(uint32_t n, const uint64_t* row_ptr, const uint32_t* col, const float* w,
                            float threshold, uint32_t maxdim, int index_words, tt_complex** out);

/* The pivot-independent part of compute_pairs' emergent-pair shortcut
 * (rips.hpp:903-928) for columns with nv vertices: per column its zero-pivot
 * cofacet (zero_pivot_cofacet_v, :585-605: the first cofacet in descending
 * index order with the column's diameter) in cof and state[t] = 1, or
 * state[t] = 2 when that cofacet has a zero-apparent facet
 * (zero_apparent_facet, :638-643), or state[t] = 0 (cof 0) when there is no
 * such cofacet. The reducer pairs a state-1 column with cof at once when cof
 * is not yet a pivot. */
tt_status tt_complex_emergent(tt_complex* cx, uint32_t nv, const uint64_t* idx, const float* d, uint64_t count,
                              uint64_t* cof, uint8_t* state);

/* The simplices (nv vertices) the next tt_complex_assemble starts from:
 * dimension 1's are the kept edges (rips.hpp:868, `simplices.swap(edges)`). */
tt_status tt_complex_seed(tt_complex* cx, uint32_t nv, const uint64_t* idx, const float* d, uint64_t count);

/* Reducer::assemble (rips.hpp:981-1023) from the seeded simplices (nv
 * vertices): every cofacet with a new top vertex up to the threshold that
 * is neither one of the npivots pivots of the reduction just done (any
 * order) nor in a zero-apparent pair either way (:1001-1007), in descending
 * filtration order (diameter descending, index ascending; :1019), with
 * each column's emergent state as tt_complex_emergent gives it. *skipped:
 * the apparent ones. keep_next: every such cofacet becomes the seeded
 * simplices (nv + 1 vertices). *idx, *cof (ncols x index_words), *d and
 * *state are malloc'd; release them with tt_free. */
tt_status tt_complex_assemble(tt_complex* cx, const uint64_t* pivots, uint64_t npivots, int keep_next,
                              uint64_t* ncols, uint64_t** idx, float** d, uint64_t** cof, uint8_t** state,
                              uint64_t* skipped);
void tt_complex_destroy(tt_complex* cx);

/* ---- edges of the boundary (RedZeD) engine -------------------------------- */

/* persistence()'s boundary branch front end (turbotops.hpp:243-246:
 * make_dense, enclosing_radius as threshold when threshold = +inf) and
 * RedZeD's build_edges over the dense complex (redzed.hpp:544-613,
 * DenseGeom::bucket_row :677-705): every edge i < j with d(i, j) <= threshold
 * in (diameter ascending, combinatorial index ascending) order, laid out as
 * EdgeList holds it (redzed.hpp:436-442): verts[2r] = i, verts[2r+1] = j,
 * diam[r]. mat (host, may be NULL) receives the matrix in `layout`. The
 * library mallocs *verts and *diam; release them with tt_free. */
tt_status tt_build_edges(const float* pts, uint32_t n, uint32_t dim, int kind, float p,
                         float threshold, int layout, float* mat, uint64_t* m, int32_t** verts,
                         float** diam, float* threshold_out);

/* tt_build_edges, and per edge the boundary engine's ready-witness flag in
 * bit 31 of its larger vertex id (verts[2e + 1]): whether a triangle {lo, hi,
 * w} with max(d(lo, w), d(hi, w)) <= d exists whose youngest face in this
 * order is the edge (w < lo; or lo < w < hi and d(hi, w) < d; or w > hi and
 * both < d). No reference counterpart: the flag lets the host engine
 * (turbotops_amd/host/zerodiff.hpp) skip its witness search for an edge with
 * no active face around it, the search RedZeD runs per edge
 * (redzed.hpp:1237-1241, has_leftover_coface). */
tt_status tt_build_edges_witness(const float* pts, uint32_t n, uint32_t dim, int kind, float p,
                                 float threshold, int layout, float* mat, uint64_t* m,
                                 int32_t** verts, float** diam, float* threshold_out);

/* build_edges over a graph (redzed.hpp:544-613 with SparseGeom, :831-979;
 * redzed::compute_graph :1452-1461): edges u < v with w <= threshold of a CSR
 * graph (rows sorted by neighbour id), in the same order and layout. */
tt_status tt_build_edges_csr(uint32_t n, const uint64_t* row_ptr, const uint32_t* col,
                             const float* w, float threshold, uint64_t* m, int32_t** verts,
                             float** diam);

/* ---- batches of clouds ---------------------------------------------------- */

/* Many clouds of uneven size and dimension through the dense front end of
 * persistence() (turbotops.hpp:230-253) in one launch sequence, recorded
 * once into a HIP graph and replayed. The reference has no batch API: this is
 * the per-cloud front end (make_dense, enclosing_radius, Reducer::dim0's sort
 * and apparent test) applied to every cloud of the batch at once.
 * Opaque handle; create once per set of shapes. threshold = +inf: each
 * cloud's enclosing radius. The tile shape is autotuned on the first run. */
typedef struct tt_batch tt_batch;
tt_status tt_batch_create(uint32_t nclouds, const uint32_t* n, const uint32_t* dim, int kind,
                          float p, float threshold, tt_batch** out);
/* pts: every cloud's points, row-major, concatenated in order. Returns once
 * the device work is queued and the per-cloud host copies are issued. */
tt_status tt_batch_run(tt_batch* b, const float* pts, int use_graph);
/* Blocks until cloud `cloud`'s results are in host memory. */
tt_status tt_batch_wait(tt_batch* b, uint32_t cloud);
/* Cloud results (library-owned, valid until the next run): the full
 * matrix, its threshold, and the kept edges that are not in an apparent
 * pair (idx, d), in filtration order — the union-find and the H1 columns
 * of dimension 0 need no others (order 0); in order 3 every kept edge. */
tt_status tt_batch_cloud(tt_batch* b, uint32_t cloud, const float** mat, float* threshold,
                         uint64_t* m, const uint64_t** idx, const float** d);
/* Which edges a run hands over. order = 0 (the default): the kept edges that
 * are not in an apparent pair, in the coboundary engine's order (d ascending,
 * index descending; rips.hpp:116-118), read with tt_batch_cloud. order = 1:
 * every kept edge in the boundary engine's order (d ascending, combinatorial
 * index ascending; redzed.hpp:544-613), read with tt_batch_cloud_edges.
 * order = 2: as 1, plus each edge's ready-witness flag (as
 * tt_build_edges_witness) where the batch's keys have a bit to spare.
 * order = 3: every kept edge in the coboundary engine's order with its
 * apparent flag (as tt_dense_edges hands them over), read with
 * tt_batch_cloud and tt_batch_cloud_apparent: the input of maxdim >= 2,
 * whose next dimension is assembled from all edges (rips.hpp:981-1023);
 * TT_ELENGTH where the keys have no bit to spare.
 * Takes effect at the next run (the recorded sequence is rebuilt). No
 * reference counterpart: the reference has no batch API. */
tt_status tt_batch_set_order(tt_batch* b, int order);
/* Order 3: the apparent flags (1 = in a zero-apparent pair) of the cloud's m
 * edges that tt_batch_cloud returns, in the same order. */
tt_status tt_batch_cloud_apparent(tt_batch* b, uint32_t cloud, const uint8_t** apparent);
/* Cloud results in orders 1 and 2: the full matrix, its threshold, and m
 * edges as 2 m vertex ids (lo < hi per edge) with their diameters. When
 * *witness_bit is 1, bit 31 of each hi id is the edge's ready-witness flag,
 * as tt_build_edges_witness hands it over (batches whose keys have no bit to
 * spare: 0, and the ids are plain). */
tt_status tt_batch_cloud_edges(tt_batch* b, uint32_t cloud, const float** mat, float* threshold,
                               uint64_t* m, const uint32_t** verts, const float** d, int* witness_bit);
/* Device time of the last run, the autotuned rows per wave (8/16/32; the
 * tile shape is tuned per dimension group, and this is the one of the group
 * with the most tiles), and the kept / non-apparent edge counts (both the
 * kept count in orders 1 and 2). */
tt_status tt_batch_info(tt_batch* b, float* device_ms, int* rows_per_wave, uint64_t* edges_total,
                        uint64_t* edges_kept);
/* The distance stage of the batch alone, re-run on the points of the last
 * run with the tuned tile shapes and timed with HIP events on the batch's
 * stream (the median of five launches after a warm-up), in milliseconds.
 * The matrices it rewrites are bit-identical. */
tt_status tt_batch_time_dist(tt_batch* b, float* ms);
void tt_batch_destroy(tt_batch* b);

/* ---- batched kNN graphs ------------------------------------------------------ */

/* The kNN stage of the batch: for every cloud, its exact k-nearest-neighbour
 * lists (knn_lists_exact, include/turbotops/graph.hpp:211-225: k_c =
 * min(k, n_c - 1), (distance, id) lexicographic, self excluded) and their
 * symmetric union as a CSR graph, rows sorted by neighbour id, weights
 * recomputed as c(row, col) (knn_graph, graph.hpp:286-304), in one launch
 * sequence recorded into a HIP graph on the first run. Each cloud's graph
 * equals tt_knn_graph_exact on that cloud alone, bit for bit. Any k >= 1: a
 * cloud whose list is longer than 64 takes the single-cloud selection path,
 * and a batch holding such a cloud runs its sequence directly (use_graph is
 * then ignored). No reference counterpart: the reference
 * has no batch API; its per-cloud path is persistence(..., graph = knn)
 * (turbotops.hpp:275-277). */
typedef struct tt_batch_knn tt_batch_knn;
tt_status tt_batch_knn_create(uint32_t nclouds, const uint32_t* n, const uint32_t* dim, int kind,
                              float p, uint32_t k, tt_batch_knn** out);
/* pts: every cloud's points, row-major, concatenated in order. Returns with
 * every cloud's graph in host memory. */
tt_status tt_batch_knn_run(tt_batch_knn* b, const float* pts, int use_graph);
/* Cloud results (library-owned, valid until the next run): n_c + 1 row
 * offsets into col / w, and col / w themselves; the offsets are batch-wide,
 * so cloud c's rows run from col + row_off[0] to col + row_off[n_c], with
 * cloud-local neighbour ids. */
tt_status tt_batch_knn_cloud(tt_batch_knn* b, uint32_t cloud, const uint32_t** row_off,
                             const uint32_t** col, const float** w);
/* Cloud c's lists, n_c x k_c ids, each row by (distance, id); copied from
 * the device per call (valid until the next call). */
tt_status tt_batch_knn_lists(tt_batch_knn* b, uint32_t cloud, uint32_t* k, const uint32_t** ids);
/* Device time of the last run and the batch's total CSR entries. */
tt_status tt_batch_knn_info(tt_batch_knn* b, float* device_ms, uint64_t* nnz);
void tt_batch_knn_destroy(tt_batch_knn* b);

/* ---- differentiation ------------------------------------------------------- */

/* Backward of a diagram with respect to the points. The reference computes
 * diagrams only (rips.hpp:52-67: Barcode keeps bar values, not simplices);
 * every finite bar value is the length of a critical edge (i, j) — the
 * longest edge of the bar's birth or death simplex — so
 *   dx[a] += grad[e] * d dist(x_a, x_b) / d x_a   for edge e = (a, b), both ends,
 * with the derivative of the metric `kind`/`p` (metric.hpp:92-207 define the
 * distances): L2 (x_a - x_b)/d (0 at d = 0); L1 sign(x_a - x_b); L-inf the sign
 * on the first coordinate of largest |x_a - x_b|; Lp sign * |diff|^(p-1)/d^(p-1).
 * edges: m pairs (2m u32, endpoints < n); grad: m floats; dx: n*dim floats,
 * overwritten. Deterministic: contributions are grouped per point by a sort
 * and summed in FP64, rounded once. */
tt_status tt_diagram_backward(const float* pts, uint32_t n, uint32_t dim, int kind, float p,
                              const uint32_t* edges, const float* grad, uint32_t m, float* dx);

/* Device form (HBM pointers, stream as hipStream_t); returns after the
 * stream has finished the backward. The endpoints are not checked on the
 * host here: an edge with an endpoint >= n contributes nothing to either end
 * (the host form rejects it with TT_EINVAL "tt_diagram_backward: edge
 * endpoint out of range"). */
tt_status tt_diagram_backward_dev(const float* d_pts, uint32_t n, uint32_t dim, int kind, float p,
                                  const uint32_t* d_edges, const float* d_grad, uint32_t m,
                                  float* d_dx, void* stream);

/* Releases a buffer the library allocated for a host-side result. */
void tt_free(void* p);

#ifdef __cplusplus
}
#endif

#endif /* TTGPU_H */
