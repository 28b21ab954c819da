/* SPDX-License-Identifier: MIT
 * ttoracle.h — CPU restatement of the turbotops filtration front end.
 *
 * TEST INFRASTRUCTURE ONLY. This is the oracle: tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg load it as the checker. The product
 * (turbotops_amd) never links, loads or calls it.
 *
 * Metric kinds match the reference's Metric::Kind order (metric.hpp:215).
 */
#ifndef TTORACLE_H
#define TTORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { TTO_L1 = 0, TTO_L2 = 1, TTO_LINF = 2, TTO_LP = 3 };

int tto_metric_parse(const char* s, int* kind, float* p);
float tto_dist(int kind, float p, const float* a, const float* b, uint32_t dim);
void tto_make_dense(const float* pts, uint32_t n, uint32_t dim, int kind, float p, int full,
                    float* out);
float tto_enclosing_radius(const float* mat, uint32_t n, int full);
uint64_t tto_radius_graph(const float* pts, uint32_t n, uint32_t dim, int kind, float p,
                          float tau, uint64_t* row_ptr, uint32_t* col, float* w);
uint32_t tto_knn(const float* pts, uint32_t n, uint32_t dim, int kind, float p, uint32_t k,
                 uint32_t* ids);
void tto_knn_rows(const float* pts, uint32_t n, uint32_t dim, int kind, float p, uint32_t k,
                  const uint32_t* rows, uint32_t nrows, uint32_t* ids);
uint64_t tto_radius_row(const float* pts, uint32_t n, uint32_t dim, int kind, float p, float tau,
                        uint32_t i, uint32_t* col, float* w);
uint64_t tto_knn_graph(const float* pts, uint32_t n, uint32_t dim, int kind, float p,
                       uint32_t k, const uint32_t* ids, uint64_t* row_ptr, uint32_t* col,
                       float* w);
void tto_mst_deaths(const float* pts, uint32_t n, uint32_t dim, int kind, float p,
                    float* deaths);
uint64_t tto_dense_edges(const float* mat, uint32_t n, float threshold, uint64_t* idx,
                         float* d);
int tto_apparent_cofacet(const float* mat, uint32_t n, uint32_t nv, uint64_t idx_lo,
                         uint64_t idx_hi, float d);
int tto_apparent_facet(const float* mat, uint32_t n, uint32_t nv, uint64_t idx_lo,
                       uint64_t idx_hi, float d);
void tto_apparent_edges(const float* mat, uint32_t n, const uint64_t* idx, const float* d,
                        uint64_t m, uint8_t* flags);
int tto_decode(uint32_t n, uint32_t nv, uint64_t idx_lo, uint64_t idx_hi, uint32_t* vd);

/* Complex-generic forms: sparse = 0 with mat (n x n full layout), or
 * sparse = 1 with a CSR graph (rp, col, w); simplex indices as `words`
 * 64-bit words each (1 or 2, low word first). */
void tto_apparent_cx(int sparse, uint32_t n, const float* mat, const uint64_t* rp, const uint32_t* col,
                     const float* w, uint32_t nv, uint32_t words, const uint64_t* idx, const float* d,
                     uint64_t count, uint8_t* out);
void tto_emergent_cx(int sparse, uint32_t n, const float* mat, const uint64_t* rp, const uint32_t* col,
                     const float* w, float threshold, uint32_t nv, uint32_t words, const uint64_t* idx,
                     const float* d, uint64_t count, uint64_t* cof, uint8_t* state);
void tto_assemble_cx(int sparse, uint32_t n, const float* mat, const uint64_t* rp, const uint32_t* col,
                     const float* w, float threshold, uint32_t nv, uint32_t words, const uint64_t* idx,
                     const float* d, uint64_t count, const uint64_t* pivots, uint64_t npivots, int keep_next,
                     uint64_t* ncols, uint64_t** cidx, float** cd, uint64_t** ccof, uint8_t** cstate,
                     uint64_t* nnext, uint64_t** nidx, float** nd, uint64_t* skipped);
void tto_free(void* p);

#ifdef __cplusplus
}
#endif

#endif
