/* SPDX-License-Identifier: MIT
 * ttoracle.c — CPU restatement of the turbotops filtration front end.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg as the checker; never by the product.
 *
 * Each function restates the reference's scalar (TURBOTOPS_NO_SIMD) code
 * path step by step, including the order of every floating-point operation,
 * and cites the reference file:line it follows. Built with
 * -ffp-contract=off, so the only fused multiply-adds are the explicit fmaf
 * calls that mirror the reference's std::fma.
 *
 * Pins (what checks each function against something other than itself):
 *   tto_dist          reference CLI goldens (unit square: 1, sqrt 2, 2, 2^(1/3)),
 *                     metric axioms, and bitwise vs the compiled reference
 *                     (oracle/_ref) on random clouds in every metric
 *   tto_make_dense /  bitwise vs oracle/_ref; lower == full; closed forms
 *   tto_enclosing_radius
 *   tto_radius_graph  CsrGraph invariants; bitwise vs oracle/_ref
 *   tto_knn           bitwise vs oracle/_ref; sorted-distance brute force
 *   tto_knn_graph     bitwise vs oracle/_ref
 *   tto_mst_deaths    Kruskal (textbook) multiset; vs oracle/_ref bars
 *   tto_dense_edges   total order (any correct sort); vs oracle/_ref
 *   tto_apparent_*    vs oracle/_ref flags; brute-force restatement in the
 *                     tests from the definition (first cofacet / facet)
 */
#include "ttoracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

/* ---- metrics (include/turbotops/metric.hpp:92-253) ---------------------- */

/* std::max(m, v) == (m < v) ? v : m; std::min(1, v) == (v < 1) ? v : 1 */
static inline float fmax_ref(float m, float v) { return (m < v) ? v : m; }

/* kern::l2, NO_SIMD branch (metric.hpp:92-112): sequential fma chain, sqrt */
static float l2(const float* a, const float* b, uint32_t dim) {
    float s = 0.0f;
    for (uint32_t k = 0; k < dim; ++k) {
        const float d = a[k] - b[k];
        s = fmaf(d, d, s);
    }
    return sqrtf(s);
}

/* kern::l1, NO_SIMD branch (metric.hpp:144-159) */
static float l1(const float* a, const float* b, uint32_t dim) {
    float s = 0.0f;
    for (uint32_t k = 0; k < dim; ++k) s += fabsf(a[k] - b[k]);
    return s;
}

/* kern::linf, NO_SIMD branch (metric.hpp:161-174) */
static float linf(const float* a, const float* b, uint32_t dim) {
    float m = 0.0f;
    for (uint32_t k = 0; k < dim; ++k) m = fmax_ref(m, fabsf(a[k] - b[k]));
    return m;
}

/* kern::ipow (metric.hpp:176-181): square-and-multiply */
static float ipow(float x, uint32_t p) {
    float r = 1.0f;
    for (; p; p >>= 1, x *= x)
        if (p & 1) r *= x;
    return r;
}

/* kern::lp, NO_SIMD branch (metric.hpp:183-207): max-normalised sum */
static float lp(const float* a, const float* b, uint32_t dim, float p) {
    const float M = linf(a, b, dim);
    if (M == 0.0f) return 0.0f;
    const float inv = 1.0f / M;
    const int integral = p == floorf(p) && p <= 8.0f;
    float s = 0.0f;
    for (uint32_t k = 0; k < dim; ++k) {
        const float v = fabsf(a[k] - b[k]) * inv;
        const float x = (v < 1.0f) ? v : 1.0f;
        s += integral ? ipow(x, (uint32_t)p) : powf(x, p);
    }
    return M * powf(s, 1.0f / p);
}

float tto_dist(int kind, float p, const float* a, const float* b, uint32_t dim) {
    switch (kind) {
        case TTO_L1: return l1(a, b, dim);
        case TTO_L2: return l2(a, b, dim);
        case TTO_LINF: return linf(a, b, dim);
        default: return lp(a, b, dim, p);
    }
}

/* Metric::minkowski / Metric::parse (metric.hpp:227-253). Returns 0 on
 * success, 1 for a rejected metric (the reference's invalid_argument). */
static int minkowski(float p, int* kind, float* pout) {
    if (!(p >= 1.0f)) return 1;
    *pout = p;
    if (p == 1.0f) *kind = TTO_L1;
    else if (p == 2.0f) *kind = TTO_L2;
    else if (isinf(p)) *kind = TTO_LINF;
    else *kind = TTO_LP;
    return 0;
}

int tto_metric_parse(const char* s, int* kind, float* p) {
    if (!strcmp(s, "l2") || !strcmp(s, "euclidean")) return minkowski(2.0f, kind, p);
    if (!strcmp(s, "l1") || !strcmp(s, "manhattan")) return minkowski(1.0f, kind, p);
    if (!strcmp(s, "linf") || !strcmp(s, "chebyshev")) return minkowski(INFINITY, kind, p);
    if (!strncmp(s, "lp:", 3)) {
        const char* v = s + 3;
        if (!strcmp(v, "inf")) return minkowski(INFINITY, kind, p);
        char* end = NULL;
        const float pv = strtof(v, &end);
        if (end && *end == '\0' && *v) return minkowski(pv, kind, p);
    }
    return 2;
}

/* ---- dense matrix (rips.hpp:262-416) ------------------------------------ */

/* make_dense: the lower triangle d(i, j), j < i, with a = x_i, b = x_j
 * (rips.hpp:398-400, :406-409), mirrored for the full layout (:401-403). */
void tto_make_dense(const float* pts, uint32_t n, uint32_t dim, int kind, float p, int full,
                    float* out) {
    if (full) {
        for (uint64_t i = 0; i < n; ++i) {
            out[i * n + i] = 0.0f;
            for (uint32_t j = 0; j < i; ++j)
                out[i * n + j] = tto_dist(kind, p, pts + i * dim, pts + (uint64_t)j * dim, dim);
        }
        for (uint64_t i = 0; i < n; ++i)
            for (uint64_t j = i + 1; j < n; ++j) out[i * n + j] = out[j * n + i];
    } else {
        for (uint64_t i = 1; i < n; ++i) {
            const uint64_t row = i * (i - 1) / 2;
            for (uint32_t j = 0; j < i; ++j)
                out[row + j] = tto_dist(kind, p, pts + i * dim, pts + (uint64_t)j * dim, dim);
        }
    }
}

static inline float dense_at(const float* mat, uint32_t n, int full, uint32_t i, uint32_t j) {
    if (i == j) return 0.0f;
    if (full) return mat[(uint64_t)i * n + j];
    if (i < j) {
        const uint32_t t = i;
        i = j;
        j = t;
    }
    return mat[(uint64_t)i * (i - 1) / 2 + j];
}

/* DenseComplex::enclosing_radius (rips.hpp:370-380): min_i max_j d(i, j),
 * j ranging over all vertices including i itself. */
float tto_enclosing_radius(const float* mat, uint32_t n, int full) {
    float best = INFINITY;
    for (uint32_t i = 0; i < n; ++i) {
        float m = 0.0f;
        for (uint32_t j = 0; j < n; ++j) m = fmax_ref(m, dense_at(mat, n, full, i, j));
        best = (m < best) ? m : best;
    }
    return best;
}

/* ---- neighbourhood graphs (graph.hpp:111-304, core.hpp:113-134) --------- */

/* radius_graph: every j != i with c(i, j) <= tau, row sorted by neighbour
 * id (CsrGraph::from_rows sorts (id, weight) pairs). Brute force replaces
 * the k-d tree, which only prunes and never changes the result. col/w may
 * be NULL to count; row_ptr (n + 1) is always filled. Returns 2m. */
uint64_t tto_radius_graph(const float* pts, uint32_t n, uint32_t dim, int kind, float p,
                          float tau, uint64_t* row_ptr, uint32_t* col, float* w) {
    uint64_t o = 0;
    row_ptr[0] = 0;
    for (uint64_t i = 0; i < n; ++i) {
        for (uint32_t j = 0; j < n; ++j) {
            if (j == i) continue;
            const float d = tto_dist(kind, p, pts + i * dim, pts + (uint64_t)j * dim, dim);
            if (d <= tau) {
                if (col) {
                    col[o] = j;
                    w[o] = d;
                }
                ++o;
            }
        }
        row_ptr[i + 1] = o;
    }
    return o;
}

typedef struct {
    float d;
    uint32_t id;
} cand_t;

/* (distance, id) lexicographic order (graph.hpp:112) */
static int cand_cmp(const void* x, const void* y) {
    const cand_t* a = (const cand_t*)x;
    const cand_t* b = (const cand_t*)y;
    if (a->d < b->d) return -1;
    if (b->d < a->d) return 1;
    return (a->id > b->id) - (a->id < b->id);
}

/* knn_lists_exact + knn_query (graph.hpp:123-166, :211-225): the k
 * smallest (c(i, j), j), j != i, ascending; k clamped to n - 1. Returns the
 * clamped k. */
uint32_t tto_knn(const float* pts, uint32_t n, uint32_t dim, int kind, float p, uint32_t k,
                 uint32_t* ids) {
    const uint32_t kk = n ? (k < n - 1 ? k : n - 1) : 0;
    if (!kk) return 0;
    cand_t* c = (cand_t*)malloc(sizeof(cand_t) * n);
    for (uint64_t i = 0; i < n; ++i) {
        uint32_t m = 0;
        for (uint32_t j = 0; j < n; ++j) {
            if (j == i) continue;
            c[m].d = tto_dist(kind, p, pts + i * dim, pts + (uint64_t)j * dim, dim);
            c[m].id = j;
            ++m;
        }
        qsort(c, m, sizeof(cand_t), cand_cmp);
        for (uint32_t t = 0; t < kk; ++t) ids[i * kk + t] = c[t].id;
    }
    free(c);
    return kk;
}

/* tto_knn for the given query rows only (knn_query per row, graph.hpp:
 * 123-166): ids holds nrows * k (k already clamped to n - 1). */
void tto_knn_rows(const float* pts, uint32_t n, uint32_t dim, int kind, float p, uint32_t k,
                  const uint32_t* rows, uint32_t nrows, uint32_t* ids) {
    cand_t* c = (cand_t*)malloc(sizeof(cand_t) * (n ? n : 1));
    for (uint32_t r = 0; r < nrows; ++r) {
        const uint64_t i = rows[r];
        uint32_t m = 0;
        for (uint32_t j = 0; j < n; ++j) {
            if (j == i) continue;
            c[m].d = tto_dist(kind, p, pts + i * dim, pts + (uint64_t)j * dim, dim);
            c[m].id = j;
            ++m;
        }
        qsort(c, m, sizeof(cand_t), cand_cmp);
        for (uint32_t t = 0; t < k; ++t) ids[(uint64_t)r * k + t] = c[t].id;
    }
    free(c);
}

/* One row of radius_graph (graph.hpp:170-201): the j != i with
 * c(i, j) <= tau, ascending, and their weights; col/w may be NULL to count.
 * Returns the row's length. */
uint64_t tto_radius_row(const float* pts, uint32_t n, uint32_t dim, int kind, float p, float tau,
                        uint32_t i, uint32_t* col, float* w) {
    uint64_t o = 0;
    for (uint32_t j = 0; j < n; ++j) {
        if (j == i) continue;
        const float d = tto_dist(kind, p, pts + (uint64_t)i * dim, pts + (uint64_t)j * dim, dim);
        if (d <= tau) {
            if (col) {
                col[o] = j;
                w[o] = d;
            }
            ++o;
        }
    }
    return o;
}

static int u32_cmp(const void* x, const void* y) {
    const uint32_t a = *(const uint32_t*)x, b = *(const uint32_t*)y;
    return (a > b) - (a < b);
}

/* knn_graph (graph.hpp:286-304): (i, j) is an edge when j is in i's list
 * or i in j's; rows sorted and deduplicated by id; weights recomputed as
 * c(v, u) for row v. ids holds n * k entries (k already clamped). col/w
 * may be NULL to count. Returns 2m. */
uint64_t tto_knn_graph(const float* pts, uint32_t n, uint32_t dim, int kind, float p,
                       uint32_t k, const uint32_t* ids, uint64_t* row_ptr, uint32_t* col,
                       float* w) {
    uint64_t* deg = (uint64_t*)calloc((uint64_t)n + 1, sizeof(uint64_t));
    for (uint64_t i = 0; i < n; ++i)
        for (uint32_t t = 0; t < k; ++t) {
            const uint32_t j = ids[i * k + t];
            if (j == i) continue;
            deg[i]++;
            deg[j]++;
        }
    uint64_t* start = (uint64_t*)calloc((uint64_t)n + 1, sizeof(uint64_t));
    for (uint32_t v = 0; v < n; ++v) start[v + 1] = start[v] + deg[v];
    uint32_t* all = (uint32_t*)malloc(sizeof(uint32_t) * (start[n] ? start[n] : 1));
    uint64_t* fill = (uint64_t*)calloc((uint64_t)n + 1, sizeof(uint64_t));
    for (uint64_t i = 0; i < n; ++i)
        for (uint32_t t = 0; t < k; ++t) {
            const uint32_t j = ids[i * k + t];
            if (j == i) continue;
            all[start[i] + fill[i]++] = j;
            all[start[j] + fill[j]++] = (uint32_t)i;
        }
    uint64_t o = 0;
    row_ptr[0] = 0;
    for (uint64_t v = 0; v < n; ++v) {
        uint32_t* r = all + start[v];
        const uint64_t len = start[v + 1] - start[v];
        qsort(r, len, sizeof(uint32_t), u32_cmp);
        for (uint64_t t = 0; t < len; ++t) {
            if (t && r[t] == r[t - 1]) continue;
            if (col) {
                col[o] = r[t];
                w[o] = tto_dist(kind, p, pts + v * dim, pts + (uint64_t)r[t] * dim, dim);
            }
            ++o;
        }
        row_ptr[v + 1] = o;
    }
    free(deg);
    free(start);
    free(all);
    free(fill);
    return o;
}

/* ---- H0 (mst.hpp:25-99) -------------------------------------------------- */

static int f32_cmp(const void* x, const void* y) {
    const float a = *(const float*)x, b = *(const float*)y;
    return (a > b) - (a < b);
}

/* Prim from vertex 0 with the (best, j) lexicographic minimum per step, then
 * the sorted deaths (mst.hpp:51-91). deaths holds n - 1 entries. */
void tto_mst_deaths(const float* pts, uint32_t n, uint32_t dim, int kind, float p,
                    float* deaths) {
    if (n < 2) return;
    float* best = (float*)malloc(sizeof(float) * n);
    uint8_t* in = (uint8_t*)calloc(n, 1);
    for (uint32_t j = 0; j < n; ++j) best[j] = INFINITY;
    uint32_t cur = 0;
    in[0] = 1;
    for (uint32_t step = 1; step < n; ++step) {
        float mb = INFINITY;
        uint32_t mj = ~0u;
        for (uint32_t j = 0; j < n; ++j) {
            if (in[j]) continue;
            const float d = tto_dist(kind, p, pts + (uint64_t)cur * dim, pts + (uint64_t)j * dim,
                                     dim);
            if (d < best[j]) best[j] = d;
            if (best[j] < mb || (best[j] == mb && j < mj)) {
                mb = best[j];
                mj = j;
            }
        }
        deaths[step - 1] = mb;
        in[mj] = 1;
        cur = mj;
    }
    qsort(deaths, n - 1, sizeof(float), f32_cmp);
    free(best);
    free(in);
}

/* ---- dimension-0 edges and apparent pairs (rips.hpp:72-118, 509-665,
 * 685-849) --------------------------------------------------------------- */

typedef struct {
    uint64_t idx;
    float d;
} edge_t;

/* precedes (rips.hpp:116-118): diameter ascending, index descending */
static int edge_cmp(const void* x, const void* y) {
    const edge_t* a = (const edge_t*)x;
    const edge_t* b = (const edge_t*)y;
    if (a->d < b->d) return -1;
    if (b->d < a->d) return 1;
    return (a->idx < b->idx) - (a->idx > b->idx);
}

/* Edges j < i of a FULL matrix with d <= threshold as (i(i-1)/2 + j, d),
 * in filtration order (rips.hpp:700-703, :782, :772-773). idx/d may be
 * NULL to count. */
uint64_t tto_dense_edges(const float* mat, uint32_t n, float threshold, uint64_t* idx,
                         float* d) {
    uint64_t m = 0;
    for (uint64_t i = 1; i < n; ++i)
        for (uint64_t j = 0; j < i; ++j) m += mat[i * n + j] <= threshold;
    if (!idx) return m;
    edge_t* e = (edge_t*)malloc(sizeof(edge_t) * (m ? m : 1));
    uint64_t o = 0;
    for (uint64_t i = 1; i < n; ++i)
        for (uint64_t j = 0; j < i; ++j) {
            const float v = mat[i * n + j];
            if (v <= threshold) {
                e[o].idx = i * (i - 1) / 2 + j;
                e[o].d = v;
                ++o;
            }
        }
    qsort(e, m, sizeof(edge_t), edge_cmp);
    for (uint64_t t = 0; t < m; ++t) {
        idx[t] = e[t].idx;
        d[t] = e[t].d;
    }
    free(e);
    return m;
}

/* C(i, k) exactly, or all-ones on overflow (the reference's Binomial throws
 * there, rips.hpp:77-87; callers stay in range). */
static u128 binom(uint32_t i, uint32_t k) {
    if (k > i) return 0;
    u128 r = 1;
    for (uint32_t t = 1; t <= k; ++t) r = r * (i - k + t) / t;
    return r;
}

#define KMAXV 32

/* decode (rips.hpp:548-566): vertices descending; v is the largest vertex
 * below the previous one with C(v, pos) <= remaining index. Linear search:
 * the reference's galloping guess only changes speed. */
static void decode(uint32_t n, uint32_t nv, u128 idx, uint32_t* vd) {
    uint32_t top = n;
    for (uint32_t i = 0; i < nv; ++i) {
        const uint32_t pos = nv - i;
        uint32_t v = top - 1;
        while (binom(v, pos) > idx) --v;
        vd[i] = v;
        idx -= binom(v, pos);
        top = v;
    }
}

int tto_decode(uint32_t n, uint32_t nv, uint64_t lo, uint64_t hi, uint32_t* vd) {
    decode(n, nv, ((u128)hi << 64) | lo, vd);
    return 0;
}

static u128 index_of(const uint32_t* vd, uint32_t nv) {
    u128 s = 0;
    for (uint32_t t = 0; t < nv; ++t) s += binom(vd[t], nv - t);
    return s;
}

/* zero_pivot_cofacet_v (rips.hpp:585-605) on a full matrix: the cofacet
 * found first when w runs from n - 1 down with cap = d, i.e. the largest w
 * outside the simplex with max_v d(v, w) <= d. Writes its vertices
 * (descending) to cv and returns 1, or returns 0. */
static int zero_pivot_cofacet(const float* mat, uint32_t n, const uint32_t* vd, uint32_t nv,
                              float d, uint32_t* cv) {
    for (int64_t wi = (int64_t)n - 1; wi >= 0; --wi) {
        const uint32_t w = (uint32_t)wi;
        int member = 0;
        float cd = d;
        for (uint32_t t = 0; t < nv; ++t) {
            if (vd[t] == w) member = 1;
            cd = fmax_ref(cd, mat[(uint64_t)vd[t] * n + w]);
        }
        if (member || !(cd <= d)) continue;
        uint32_t o = 0, t = 0;
        while (t < nv && vd[t] > w) cv[o++] = vd[t++];
        cv[o++] = w;
        while (t < nv) cv[o++] = vd[t++];
        return 1;
    }
    return 0;
}

/* zero_pivot_facet_v (rips.hpp:609-636): the first facet in ascending
 * index order (vertex j = 0, 1, ... removed) that keeps an edge of length
 * exactly d. Writes its vertices to fv and returns 1, or returns 0. */
static int zero_pivot_facet(const float* mat, uint32_t n, const uint32_t* vd, uint32_t nv,
                            float d, uint32_t* fv) {
    if (nv < 2) return 0;
    for (uint32_t j = 0; j < nv; ++j) {
        int hit = 0;
        for (uint32_t a = 0; a < nv && !hit; ++a)
            for (uint32_t b = a + 1; b < nv && !hit; ++b)
                if (a != j && b != j && mat[(uint64_t)vd[a] * n + vd[b]] == d) hit = 1;
        if (hit) {
            for (uint32_t i = 0, o = 0; i < nv; ++i)
                if (i != j) fv[o++] = vd[i];
            return 1;
        }
    }
    return 0;
}

/* zero_apparent_cofacet (rips.hpp:644-649, :655-659): the simplex's zero
 * pivot cofacet exists and that cofacet's zero pivot facet is the simplex. */
int tto_apparent_cofacet(const float* mat, uint32_t n, uint32_t nv, uint64_t lo, uint64_t hi,
                         float d) {
    uint32_t vd[KMAXV], cv[KMAXV + 1], fv[KMAXV + 1];
    const u128 idx = ((u128)hi << 64) | lo;
    decode(n, nv, idx, vd);
    if (!zero_pivot_cofacet(mat, n, vd, nv, d, cv)) return 0;
    if (!zero_pivot_facet(mat, n, cv, nv + 1, d, fv)) return 0;
    return index_of(fv, nv) == idx;
}

/* zero_apparent_facet (rips.hpp:638-643, :650-654) */
int tto_apparent_facet(const float* mat, uint32_t n, uint32_t nv, uint64_t lo, uint64_t hi,
                       float d) {
    uint32_t vd[KMAXV], fv[KMAXV], cv[KMAXV + 1];
    const u128 idx = ((u128)hi << 64) | lo;
    decode(n, nv, idx, vd);
    if (!zero_pivot_facet(mat, n, vd, nv, d, fv)) return 0;
    if (!zero_pivot_cofacet(mat, n, fv, nv - 1, d, cv)) return 0;
    return index_of(cv, nv) == idx;
}

/* The dim-0 apparent test per edge (rips.hpp:844-849) on a full matrix. */
void tto_apparent_edges(const float* mat, uint32_t n, const uint64_t* idx, const float* d,
                        uint64_t m, uint8_t* flags) {
    for (uint64_t t = 0; t < m; ++t) flags[t] = (uint8_t)tto_apparent_cofacet(mat, n, 2, idx[t], 0, d[t]);
}

/* ---- the reducer's per-simplex tests on a dense or sparse complex -------
 * A complex is the full matrix (rips.hpp:262-381, DenseComplex with
 * full = true) or a CSR graph's flag complex (SparseComplex, rips.hpp:
 * 418-472), with the filtration threshold. Simplex indices are passed as
 * `words` 64-bit words each (1: u64, 2: u128 low word first). */

typedef struct {
    int sparse;
    uint32_t n;
    const float* mat;        /* dense: n x n, full layout */
    const uint64_t* rp;      /* sparse: CSR rows sorted by neighbour id */
    const uint32_t* col;
    const float* w;
    float threshold;
} cx_t;

/* DenseComplex::dist (rips.hpp:267-273) / SparseComplex::dist (:428):
 * 0 on the diagonal; a missing graph edge is +inf (CsrGraph::weight,
 * core.hpp:100-105). */
static int cx_edge(const cx_t* c, uint32_t i, uint32_t j, float* out) {
    if (i == j) {
        *out = 0.0f;
        return 1;
    }
    if (!c->sparse) {
        *out = c->mat[(uint64_t)i * c->n + j];
        return 1;
    }
    uint64_t lo = c->rp[i], hi = c->rp[i + 1];
    while (lo < hi) {
        const uint64_t mid = lo + (hi - lo) / 2;
        if (c->col[mid] < j) lo = mid + 1;
        else hi = mid;
    }
    if (lo < c->rp[i + 1] && c->col[lo] == j) {
        *out = c->w[lo];
        return 1;
    }
    *out = INFINITY;
    return 0;
}

static float cx_dist(const cx_t* c, uint32_t i, uint32_t j) {
    float v;
    cx_edge(c, i, j, &v);
    return v;
}

/* Cofacet enumeration of a simplex (vd: k vertices, descending; diameter
 * diam), in descending w (rips.hpp:326-343, cofacets_full's NO_SIMD loop,
 * and SparseComplex::cofacets :432-471, the common neighbours of the rows):
 * w outside the simplex with cd = max(diam, d(v, w) for v in vd) <= cap;
 * all == 0 keeps only w > vd[0]. Calls emit(ctx, w, t, cd), t = number of
 * simplex vertices above w, until it returns 0. */
typedef int (*emit_fn)(void* ctx, uint32_t w, uint32_t t, float cd);

static void cx_cofacets(const cx_t* c, const uint32_t* vd, uint32_t k, float diam, int all, float cap,
                        emit_fn emit, void* ctx) {
    const uint32_t stop = all ? 0 : vd[0] + 1;
    uint32_t t = 0;
    for (int64_t wi = (int64_t)c->n - 1; wi >= (int64_t)stop; --wi) {
        const uint32_t w = (uint32_t)wi;
        while (t < k && vd[t] > w) ++t;
        if (t < k && vd[t] == w) continue;
        float cd = diam;
        int ok = 1;
        for (uint32_t i = 0; i < k && ok; ++i) {
            float dv;
            if (!cx_edge(c, vd[i], w, &dv)) ok = 0;   /* not a common neighbour */
            else cd = fmax_ref(cd, dv);
        }
        if (ok && cd <= cap && !emit(ctx, w, t, cd)) return;
    }
}

/* cofacet of vd (k vertices) with w inserted below the t vertices above it */
static u128 cofacet_index(const uint32_t* vd, uint32_t k, uint32_t w, uint32_t t) {
    u128 s = 0;
    for (uint32_t i = 0; i < t; ++i) s += binom(vd[i], k + 1 - i);
    s += binom(w, k + 1 - t);
    for (uint32_t i = t; i < k; ++i) s += binom(vd[i], k - i);
    return s;
}

typedef struct {
    int found;
    uint32_t w, t;
    float cd;
} first_t;

static int take_first(void* ctx, uint32_t w, uint32_t t, float cd) {
    first_t* f = (first_t*)ctx;
    f->found = 1;
    f->w = w;
    f->t = t;
    f->cd = cd;
    return 0;
}

/* zero_pivot_cofacet_v (rips.hpp:585-605): the first cofacet (descending
 * index) with cap = d; its vertices to cv. */
static int cx_zero_pivot_cofacet(const cx_t* c, const uint32_t* vd, uint32_t k, float d, uint32_t* cv,
                                 u128* idx) {
    first_t f = {0, 0, 0, 0.0f};
    cx_cofacets(c, vd, k, d, 1, d, take_first, &f);
    if (!f.found) return 0;
    for (uint32_t i = 0; i < f.t; ++i) cv[i] = vd[i];
    cv[f.t] = f.w;
    for (uint32_t i = f.t; i < k; ++i) cv[i + 1] = vd[i];
    *idx = cofacet_index(vd, k, f.w, f.t);
    return 1;
}

/* zero_pivot_facet_v (rips.hpp:609-636): the first facet in ascending index
 * order (vertex j = 0, 1, ... removed) keeping an edge of length d. */
static int cx_zero_pivot_facet(const cx_t* c, const uint32_t* vd, uint32_t k, float d, uint32_t* fv, u128* idx) {
    if (k < 2) return 0;
    for (uint32_t j = 0; j < k; ++j) {
        int hit = 0;
        for (uint32_t a = 0; a < k && !hit; ++a)
            for (uint32_t b = a + 1; b < k && !hit; ++b)
                if (a != j && b != j && cx_dist(c, vd[a], vd[b]) == d) hit = 1;
        if (hit) {
            uint32_t o = 0;
            for (uint32_t i = 0; i < k; ++i)
                if (i != j) fv[o++] = vd[i];
            *idx = index_of(fv, k - 1);
            return 1;
        }
    }
    return 0;
}

/* zero_apparent_cofacet_v (rips.hpp:644-649) */
static int cx_apparent_cofacet(const cx_t* c, const uint32_t* vd, uint32_t k, u128 idx, float d) {
    uint32_t cv[KMAXV + 2], fv[KMAXV + 2];
    u128 ci, fi;
    if (!cx_zero_pivot_cofacet(c, vd, k, d, cv, &ci)) return 0;
    if (!cx_zero_pivot_facet(c, cv, k + 1, d, fv, &fi)) return 0;
    return fi == idx;
}

/* zero_apparent_facet_v (rips.hpp:638-643) */
static int cx_apparent_facet(const cx_t* c, const uint32_t* vd, uint32_t k, u128 idx, float d) {
    uint32_t cv[KMAXV + 2], fv[KMAXV + 2];
    u128 ci, fi;
    if (!cx_zero_pivot_facet(c, vd, k, d, fv, &fi)) return 0;
    if (!cx_zero_pivot_cofacet(c, fv, k - 1, d, cv, &ci)) return 0;
    return ci == idx;
}

static cx_t make_cx(int sparse, uint32_t n, const float* mat, const uint64_t* rp, const uint32_t* col,
                    const float* w, float threshold) {
    cx_t c;
    c.sparse = sparse;
    c.n = n;
    c.mat = mat;
    c.rp = rp;
    c.col = col;
    c.w = w;
    c.threshold = threshold;
    return c;
}

static u128 get_index(const uint64_t* words, uint32_t nw, uint64_t t) {
    return nw == 2 ? (((u128)words[2 * t + 1]) << 64) | words[2 * t] : (u128)words[t];
}

static void put_index(uint64_t* words, uint32_t nw, uint64_t t, u128 v) {
    if (nw == 2) {
        words[2 * t] = (uint64_t)v;
        words[2 * t + 1] = (uint64_t)(v >> 64);
    } else {
        words[t] = (uint64_t)v;
    }
}

/* Both apparent tests of simplices with nv vertices: bit 0 of out[t]
 * zero_apparent_cofacet, bit 1 zero_apparent_facet (rips.hpp:638-665). */
void tto_apparent_cx(int sparse, uint32_t n, const float* mat, const uint64_t* rp, const uint32_t* col,
                     const float* w, uint32_t nv, uint32_t words, const uint64_t* idx, const float* d,
                     uint64_t count, uint8_t* out) {
    const cx_t c = make_cx(sparse, n, mat, rp, col, w, INFINITY);
    uint32_t vd[KMAXV + 2];
    for (uint64_t t = 0; t < count; ++t) {
        const u128 s = get_index(idx, words, t);
        decode(n, nv, s, vd);
        out[t] = (uint8_t)(cx_apparent_cofacet(&c, vd, nv, s, d[t]) | (cx_apparent_facet(&c, vd, nv, s, d[t]) << 1));
    }
}

/* The pivot-independent part of compute_pairs' emergent shortcut
 * (rips.hpp:903-928) per column (nv vertices): the first cofacet of equal
 * diameter in descending index order, cof[t] (state 1), unless that cofacet
 * has a zero-apparent facet (state 2); state 0 when there is none. The
 * column is emergent iff state == 1 and cof[t] is not yet a pivot. */
static uint8_t emergent_of(const cx_t* c, const uint32_t* vd, uint32_t nv, float d, u128* cof) {
    uint32_t cv[KMAXV + 2];
    u128 ci;
    if (!cx_zero_pivot_cofacet(c, vd, nv, d, cv, &ci)) return 0;
    *cof = ci;
    return cx_apparent_facet(c, cv, nv + 1, ci, d) ? 2 : 1;
}

void tto_emergent_cx(int sparse, uint32_t n, const float* mat, const uint64_t* rp, const uint32_t* col,
                     const float* w, float threshold, uint32_t nv, uint32_t words, const uint64_t* idx,
                     const float* d, uint64_t count, uint64_t* cof, uint8_t* state) {
    const cx_t c = make_cx(sparse, n, mat, rp, col, w, threshold);
    uint32_t vd[KMAXV + 2];
    for (uint64_t t = 0; t < count; ++t) {
        const u128 s = get_index(idx, words, t);
        decode(n, nv, s, vd);
        u128 ci = 0;
        state[t] = emergent_of(&c, vd, nv, d[t], &ci);
        put_index(cof, words, t, state[t] ? ci : 0);
    }
}

typedef struct {
    u128 idx;
    float d;
} simplex_t;

/* descending filtration order (sort_desc, rips.hpp:776-778 with precedes
 * :116-118 reversed): diameter descending, index ascending */
static int desc_cmp(const void* x, const void* y) {
    const simplex_t* a = (const simplex_t*)x;
    const simplex_t* b = (const simplex_t*)y;
    if (a->d != b->d) return a->d > b->d ? -1 : 1;
    return a->idx < b->idx ? -1 : (a->idx > b->idx ? 1 : 0);
}

static int sorted_has(const uint64_t* keys, uint32_t words, uint64_t np, u128 v) {
    uint64_t lo = 0, hi = np;
    while (lo < hi) {
        const uint64_t mid = lo + (hi - lo) / 2;
        if (get_index(keys, words, mid) < v) lo = mid + 1;
        else hi = mid;
    }
    return lo < np && get_index(keys, words, lo) == v;
}

typedef struct {
    const cx_t* c;
    const uint32_t* vd;
    uint32_t nv;
    const uint64_t* piv;
    uint32_t words;
    uint64_t np;
    int keep;
    simplex_t *cols, *next;
    uint64_t ncols, nnext, cap_cols, cap_next, skipped;
} asm_t;

static void push_simplex(simplex_t** buf, uint64_t* len, uint64_t* cap, u128 idx, float d) {
    if (*len == *cap) {
        *cap = *cap ? 2 * *cap : 1024;
        *buf = (simplex_t*)realloc(*buf, *cap * sizeof(simplex_t));
    }
    (*buf)[*len].idx = idx;
    (*buf)[*len].d = d;
    ++*len;
}

static int assemble_emit(void* ctx, uint32_t w, uint32_t t, float cd) {
    asm_t* a = (asm_t*)ctx;
    const u128 cf = cofacet_index(a->vd, a->nv, w, t);
    if (a->keep) push_simplex(&a->next, &a->nnext, &a->cap_next, cf, cd);
    if (sorted_has(a->piv, a->words, a->np, cf)) return 1;   /* cleared */
    uint32_t cv[KMAXV + 2];
    for (uint32_t i = 0; i < t; ++i) cv[i] = a->vd[i];
    cv[t] = w;
    for (uint32_t i = t; i < a->nv; ++i) cv[i + 1] = a->vd[i];
    if (cx_apparent_cofacet(a->c, cv, a->nv + 1, cf, cd) || cx_apparent_facet(a->c, cv, a->nv + 1, cf, cd)) {
        ++a->skipped;
        return 1;
    }
    push_simplex(&a->cols, &a->ncols, &a->cap_cols, cf, cd);
    return 1;
}

/* Reducer::assemble (rips.hpp:981-1023): the columns of dimension nv (nv +
 * 1 vertices) from the simplices with nv vertices — every cofacet with a
 * new top vertex (all = false) up to the threshold that is neither a pivot
 * of the reduction just done (cleared; pivots sorted ascending) nor in a
 * zero-apparent pair either way — in descending filtration order, with
 * each column's emergent state (as tto_emergent_cx). keep_next: every such
 * cofacet, in enumeration order, becomes the next simplices. Outputs are
 * malloc'd (tto_free). */
void tto_assemble_cx(int sparse, uint32_t n, const float* mat, const uint64_t* rp, const uint32_t* col,
                     const float* w, float threshold, uint32_t nv, uint32_t words, const uint64_t* idx,
                     const float* d, uint64_t count, const uint64_t* pivots, uint64_t npivots, int keep_next,
                     uint64_t* ncols, uint64_t** cidx, float** cd, uint64_t** ccof, uint8_t** cstate,
                     uint64_t* nnext, uint64_t** nidx, float** nd, uint64_t* skipped) {
    const cx_t c = make_cx(sparse, n, mat, rp, col, w, threshold);
    asm_t a;
    memset(&a, 0, sizeof(a));
    a.c = &c;
    a.nv = nv;
    a.piv = pivots;
    a.words = words;
    a.np = npivots;
    a.keep = keep_next;
    uint32_t vd[KMAXV + 2];
    for (uint64_t t = 0; t < count; ++t) {
        decode(n, nv, get_index(idx, words, t), vd);
        a.vd = vd;
        cx_cofacets(&c, vd, nv, d[t], 0, threshold, assemble_emit, &a);
    }
    qsort(a.cols, a.ncols, sizeof(simplex_t), desc_cmp);
    *ncols = a.ncols;
    *cidx = (uint64_t*)malloc((a.ncols ? a.ncols : 1) * words * 8);
    *ccof = (uint64_t*)malloc((a.ncols ? a.ncols : 1) * words * 8);
    *cd = (float*)malloc((a.ncols ? a.ncols : 1) * 4);
    *cstate = (uint8_t*)malloc(a.ncols ? a.ncols : 1);
    for (uint64_t t = 0; t < a.ncols; ++t) {
        put_index(*cidx, words, t, a.cols[t].idx);
        (*cd)[t] = a.cols[t].d;
        decode(n, nv + 1, a.cols[t].idx, vd);
        u128 ci = 0;
        (*cstate)[t] = emergent_of(&c, vd, nv + 1, a.cols[t].d, &ci);
        put_index(*ccof, words, t, (*cstate)[t] ? ci : 0);
    }
    *nnext = a.nnext;
    *nidx = (uint64_t*)malloc((a.nnext ? a.nnext : 1) * words * 8);
    *nd = (float*)malloc((a.nnext ? a.nnext : 1) * 4);
    for (uint64_t t = 0; t < a.nnext; ++t) {
        put_index(*nidx, words, t, a.next[t].idx);
        (*nd)[t] = a.next[t].d;
    }
    *skipped = a.skipped;
    free(a.cols);
    free(a.next);
}

void tto_free(void* p) { free(p); }
