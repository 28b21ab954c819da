// SPDX-License-Identifier: MIT
// knn.hpp — one query's exact k nearest neighbours, one query per lane,
// shared by the single-cloud scan (graph.hip) and the batched one
// (batch_knn.hip). The reference's rule (knn_query, graph.hpp:123-166): the
// k smallest (distance, id) pairs in lexicographic order over j != query,
// k = min(k, n - 1) (graph.hpp:214).
#pragma once

#include "pairs.hpp"

namespace ttg {

// The k best (distance, id) pairs so far, sorted, in VGPRs: an insertion
// network over KMAX slots (slots past k hold -inf and never change).
// Candidates may arrive in any id order: admission compares (distance, id)
// with the current k-th, as the reference's lexicographic rule does.
template <int KMAX>
struct KnnList {
    float bd[KMAX];
    u32 bid[KMAX];
    float kth;    // the current k-th (distance, id); id ~0 while slots are free
    u32 kthj;

    __device__ __forceinline__ void init(u32 k) {
#pragma unroll
        for (int t = 0; t < KMAX; ++t) {
            bd[t] = u32(t) < k ? INFINITY : -INFINITY;
            bid[t] = 0xffffffffu;
        }
        kth = INFINITY;
        kthj = 0xffffffffu;
    }
    __device__ __forceinline__ bool admits(float d, u32 j) const { return d < kth || (d == kth && j < kthj); }
    __device__ __forceinline__ void insert(float d, u32 j, u32 k) {
        float cd = d;
        u32 cj = j;
#pragma unroll
        for (int t = 0; t < KMAX; ++t) {
            const bool sw = cd < bd[t] || (cd == bd[t] && cj < bid[t]);
            const float td = bd[t];
            const u32 tj = bid[t];
            bd[t] = sw ? cd : td;
            bid[t] = sw ? cj : tj;
            cd = sw ? td : cd;
            cj = sw ? tj : cj;
        }
        float nk = bd[0];
        u32 nj = bid[0];
#pragma unroll
        for (int t = 1; t < KMAX; ++t)
            if (u32(t) == k - 1) {
                nk = bd[t];
                nj = bid[t];
            }
        kth = nk;
        kthj = nj;
    }
};

// Candidates j in [jlo, jhi) of the points x (row-major, DIM floats) against
// the query q (its own index qi is skipped), compile-time dimension. The
// candidate is the same for every lane, so its coordinates come through
// scalar loads and feed the VALU as SGPR operands. The ordering key
// (pdist<KEY>: for L2 the sum before the square root) rejects most candidates
// without the exact square root: only keys at or below key_reject_above(k-th
// distance) become distances and meet the (distance, id) rule. Four
// candidates per step: their keys first, then one test whether any can enter.
template <int V, int DIM, int KMAX>
__device__ __forceinline__ void knn_scan_static(const float* __restrict__ x, u32 jlo, u32 jhi, u32 qi,
                                                const float (&q)[DIM], const MetricArgs& ma, u32 k,
                                                KnnList<KMAX>& L) {
    float reject = key_reject_above<V>(L.kth);   // keys above this cannot enter
    auto admit = [&](u32 j, float key) {
        if (key <= reject && j != qi) {
            const float d = key_to_dist<V>(key);
            if (L.admits(d, j)) {
                L.insert(d, j, k);
                reject = key_reject_above<V>(L.kth);
            }
        }
    };
    u32 j = jlo;
    for (; j + 4 <= jhi; j += 4) {
        float key[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            float c[DIM];
#pragma unroll
            for (int t = 0; t < DIM; ++t) c[t] = x[u64(j + u) * DIM + t];
            key[u] = pdist<V, DIM, true>(c, q, DIM, ma);
        }
        const float kmin = __builtin_fminf(__builtin_fminf(key[0], key[1]), __builtin_fminf(key[2], key[3]));
        if (kmin <= reject) {
#pragma unroll
            for (int u = 0; u < 4; ++u) admit(j + u, key[u]);
        }
    }
    for (; j < jhi; ++j) {
        float c[DIM];
#pragma unroll
        for (int t = 0; t < DIM; ++t) c[t] = x[u64(j) * DIM + t];
        admit(j, pdist<V, DIM, true>(c, q, DIM, ma));
    }
}

// Candidates at sorted positions [jlo, jhi) of xs (points in a spatial
// order; perm[j] is the original id of position j) against the query at
// position qp, compile-time dimension. Only candidates with distance <= cap
// are admitted, where cap is at least the query's true k-th distance (the
// k-th distance within some subset of the candidates): every member of the
// true k best passes, so the list this leaves is still the k best of the
// range among them, while most insertions (which a whole wave pays for when
// any lane makes one) are avoided. Ids and ties are the original ids.
template <int V, int DIM, int KMAX>
__device__ __forceinline__ void knn_scan_sorted(const float* __restrict__ xs, const u32* __restrict__ perm,
                                                u32 jlo, u32 jhi, u32 qp, const float (&q)[DIM],
                                                const MetricArgs& ma, u32 k, float cap, KnnList<KMAX>& L) {
    float reject = key_reject_above<V>(cap);
    auto admit = [&](u32 j, float key) {
        if (key <= reject && j != qp) {
            const float d = key_to_dist<V>(key);
            const u32 id = perm[j];
            if (d <= cap && L.admits(d, id)) {
                L.insert(d, id, k);
                reject = key_reject_above<V>(L.kth < cap ? L.kth : cap);
            }
        }
    };
    // four candidates per step: their keys first, then one test whether any
    // can enter (most cannot), then each in order
    u32 j = jlo;
    for (; j + 4 <= jhi; j += 4) {
        float key[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            float c[DIM];
#pragma unroll
            for (int t = 0; t < DIM; ++t) c[t] = xs[u64(j + u) * DIM + t];
            key[u] = pdist<V, DIM, true>(c, q, DIM, ma);
        }
        const float kmin = __builtin_fminf(__builtin_fminf(key[0], key[1]), __builtin_fminf(key[2], key[3]));
        if (kmin <= reject) {
#pragma unroll
            for (int u = 0; u < 4; ++u) admit(j + u, key[u]);
        }
    }
    for (; j < jhi; ++j) {
        float c[DIM];
#pragma unroll
        for (int t = 0; t < DIM; ++t) c[t] = xs[u64(j) * DIM + t];
        admit(j, pdist<V, DIM, true>(c, q, DIM, ma));
    }
}

// The same for a runtime dimension: query and candidate read from memory
// (the query's row stays in the L1 cache; the candidate is one address per
// wave).
template <int V, int KMAX>
__device__ __forceinline__ void knn_scan_runtime(const float* __restrict__ x, u32 dim, u32 jlo, u32 jhi, u32 qi,
                                                 const MetricArgs& ma, u32 k, KnnList<KMAX>& L) {
    const float* __restrict__ qp = x + u64(qi) * dim;
    for (u32 j = jlo; j < jhi; ++j) {
        const float d = pdist<V, 0>(qp, x + u64(j) * dim, dim, ma);
        if (j != qi && L.admits(d, j)) L.insert(d, j, k);
    }
}

}  // namespace ttg
