// SPDX-License-Identifier: MIT
// graph.hip — exact neighbourhood graphs over a point cloud.
//
// Replaces radius_graph (include/turbotops/graph.hpp:170-201),
// knn_lists_exact / knn_query (graph.hpp:123-166, :211-225), knn_graph
// (graph.hpp:286-304) and CsrGraph::from_rows (core.hpp:113-134). The
// reference prunes with a k-d tree; pruning never changes its result, so an
// exhaustive tiled scan reproduces it exactly.
//
//   radius graph  64x64 tiles of the lower triangle (as the distance matrix)
//                 over the points in a spatial order; a tile whose blocks' boxes
//                 lie farther apart than tau is skipped whole (a lower bound of
//                 every pair's computed distance); per wave and row a ballot of
//                 d(i, j) <= tau, one atomic per ballot reserves room in an edge
//                 list, mbcnt places each edge.
//                 The CSR comes from a radix sort of (row, col) keys carrying the
//                 weight bits: rows sorted by neighbour id, as from_rows.
//   kNN           one query per lane (knn.hpp); candidates stream through
//                 scalar loads; the k best (distance, id) pairs sit in VGPRs as
//                 a sorted insertion network, admitting by the reference's
//                 (distance, id) lexicographic rule (graph.hpp:112, :145).
//                 Every pair is measured; queries and candidates follow a
//                 spatial order, and a bound from each query's neighbourhood
//                 in that order keeps far candidates out of the lists.
//   kNN, k > 64   one workgroup per query: a radix select over the 32-bit
//                 distance bits (four 8-bit passes, recomputing the distances)
//                 finds the k-th distance D*, a second select over the ids
//                 breaks the ties at D* (only when they do not all fit); the k
//                 chosen (distance, id) pairs are then ordered by two stable
//                 radix sorts over every query at once.
//   kNN graph     both directions of every list entry as (row, col) keys, radix
//                 sort, adjacent-duplicate removal, weights recomputed as
//                 c(row, col) (graph.hpp:295-302).
#include <algorithm>

#include "frontend.hpp"
#include "knn.hpp"
#include "pairs.hpp"

namespace ttg {

namespace {

constexpr u32 GT = 64;       // tile edge
constexpr u32 GB = 256;      // threads

__device__ __forceinline__ void tile_of_g(u64 b, u32& bi, u32& bj) {
    u32 t = u32((__builtin_sqrt(8.0 * double(b) + 1.0) - 1.0) * 0.5);
    while (u64(t) * (t + 1) / 2 > b) --t;
    while (u64(t + 1) * (t + 2) / 2 <= b) ++t;
    bi = t;
    bj = u32(b - u64(t) * (t + 1) / 2);
}

__device__ __forceinline__ u32 mbcnt64(u64 m) {
    return __builtin_amdgcn_mbcnt_hi(u32(m >> 32), __builtin_amdgcn_mbcnt_lo(u32(m), 0u));
}

// Key between the nearest points of two boxes (coordinate minima then
// maxima): per coordinate the gap between the intervals, combined as the
// metric's key (the fma sum of squares for L2, the sum for L1, the largest
// gap for L-inf and Lp). For points a, b in the two boxes each gap is at most
// |fl(a_k - b_k)| (the subtractions round monotonically), and the sums, fused
// multiply-adds and maxima round monotonically too, so this is a lower bound
// of the pair's key as computed; for Lp the largest coordinate difference m is
// one (m * pow(s, 1/p) with s >= 1 is never below m).
template <int V, int DIM>
__device__ __forceinline__ float box_gap_key(const float* __restrict__ a, const float* __restrict__ b, u32 dim_rt) {
    const u32 dim = DIM > 0 ? u32(DIM) : dim_rt;
    float s = 0.0f;
#pragma unroll
    for (u32 k = 0; k < dim; ++k) {
        const float g = __builtin_fmaxf(0.0f, __builtin_fmaxf(a[k] - b[dim + k], b[k] - a[dim + k]));
        if constexpr (V == MV_L2) s = __builtin_fmaf(g, g, s);
        else if constexpr (V == MV_L1) s += g;
        else s = __builtin_fmaxf(s, g);
    }
    return s;
}

// Edge list of every pair with d <= tau, over the points in spatial order
// (xs; perm[pos] = original id): 64x64 tiles of the lower triangle of
// positions, each workgroup looping over tiles. A tile whose two blocks'
// boxes are farther apart than tau holds no edge and is skipped whole;
// otherwise every pair j < i of it is measured. Edges carry original ids;
// entries past `cap` are counted but not stored.
template <int V, int DIM>
__global__ __launch_bounds__(GB) void radius_tiles(const float* __restrict__ x, const u32* __restrict__ perm,
                                                   const float* __restrict__ box, const float* __restrict__ box16,
                                                   u32 n, u32 dim_rt,
                                                   MetricArgs ma, float tau, u64 tiles, u64 cap,
                                                   unsigned long long* __restrict__ counter,
                                                   u32* __restrict__ ei, u32* __restrict__ ej,
                                                   float* __restrict__ ew) {
    const u32 dim = DIM > 0 ? u32(DIM) : dim_rt;
    const u32 lane = threadIdx.x & 63;
    const u32 wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    // Compile-time dimensions measure ordering keys (for L2 the sum before
    // the square root): a key above key_reject_above(tau) cannot pass, so only
    // the rest take the exact square root; the test itself stays d <= tau on
    // the distance.
    const float reject = DIM > 0 ? key_reject_above<V>(tau) : INFINITY;
    for (u64 t = blockIdx.x; t < tiles; t += gridDim.x) {
        u32 bi, bj;
        tile_of_g(t, bi, bj);
        const float* bxj = box + u64(bj) * 2 * dim;
        if (key_to_dist<V>(box_gap_key<V, DIM>(box + u64(bi) * 2 * dim, bxj, dim)) > tau)
            continue;   // block-uniform
        const u32 j = bj * GT + lane;
        const u32 jc = j < n ? j : n - 1;
        const u32 ibase = bi * GT + wave * ROWS;
        // the same test for this wave's 16 rows alone (the waves share no LDS
        // and no barrier, so each skips on its own)
        if (ibase >= n ||
            key_to_dist<V>(box_gap_key<V, DIM>(box16 + u64(ibase / ROWS) * 2 * dim, bxj, dim)) > tau)
            continue;   // wave-uniform
        float val[ROWS];
        if constexpr (DIM > 0) rows_static<V, DIM, ROWS, true>(x, n, ma, ibase, jc, val);
        else rows_runtime<V>(x, n, dim_rt, ma, ibase, jc, val);
#pragma unroll
        for (u32 r = 0; r < ROWS; ++r) {
            const u32 i = ibase + r;
            bool keep = i < n && j < i && val[r] <= reject;
            if constexpr (DIM > 0) {
                if (keep) val[r] = key_to_dist<V>(val[r]);
            }
            keep = keep && val[r] <= tau;
            const u64 b = __ballot(keep);
            if (b) {
                unsigned long long base = 0;
                if (lane == 0) base = atomicAdd(counter, (unsigned long long)__popcll(b));
                base = __shfl(base, 0);
                if (keep) {
                    const u64 pos = base + mbcnt64(b);
                    if (pos < cap) {
                        ei[pos] = perm[i];
                        ej[pos] = perm[j];
                        ew[pos] = val[r];
                    }
                }
            }
        }
    }
}

template <int V, int DIM>
struct RadiusLaunch {
    static void run(u32 blocks, hipStream_t s, const float* x, const u32* perm, const float* box,
                    const float* box16, u32 n, u32 dim, MetricArgs ma, float tau, u64 tiles, u64 cap,
                    unsigned long long* counter, u32* ei, u32* ej, float* ew) {
        radius_tiles<V, DIM><<<blocks, GB, 0, s>>>(x, perm, box, box16, n, dim, ma, tau, tiles, cap, counter, ei,
                                                   ej, ew);
    }
};

// Both directions of every edge as (row << cb | col) keys with the weight bits.
__global__ void emit_both(const u32* __restrict__ ei, const u32* __restrict__ ej,
                          const float* __restrict__ ew, u64 m, int cb, u64* __restrict__ keys,
                          u32* __restrict__ vals, u32* __restrict__ deg) {
    for (u64 t = blockIdx.x * u64(blockDim.x) + threadIdx.x; t < m; t += u64(gridDim.x) * blockDim.x) {
        const u32 a = ei[t], b = ej[t];
        const u32 w = __float_as_uint(ew[t]);
        keys[2 * t] = (u64(a) << cb) | b;
        vals[2 * t] = w;
        keys[2 * t + 1] = (u64(b) << cb) | a;
        vals[2 * t + 1] = w;
        atomicAdd(deg + a, 1u);
        atomicAdd(deg + b, 1u);
    }
}

__global__ void split_keys(const u64* __restrict__ keys, const u32* __restrict__ vals, u64 m,
                           u64 mask, u32* __restrict__ col, float* __restrict__ w) {
    for (u64 t = blockIdx.x * u64(blockDim.x) + threadIdx.x; t < m; t += u64(gridDim.x) * blockDim.x) {
        col[t] = u32(keys[t] & mask);
        w[t] = __uint_as_float(vals[t]);
    }
}

__global__ void widen_u32(const u32* __restrict__ in, u64* __restrict__ out, u64 count) {
    for (u64 t = blockIdx.x * u64(blockDim.x) + threadIdx.x; t < count; t += u64(gridDim.x) * blockDim.x)
        out[t] = in[t];
}

// ---- kNN --------------------------------------------------------------------
// Runtime dimension: one query per lane over every candidate (knn.hpp).
template <int V, int KMAX>
__global__ __launch_bounds__(GB) void knn_scan(const float* __restrict__ x, u32 n, u32 dim, MetricArgs ma, u32 k,
                                               u32* __restrict__ ids) {
    const u32 i = blockIdx.x * GB + threadIdx.x;
    const u32 qi = i < n ? i : n - 1;
    KnnList<KMAX> L;
    L.init(k);
    knn_scan_runtime<V, KMAX>(x, dim, 0, n, qi, ma, k, L);
    if (i < n) {
#pragma unroll
        for (int t = 0; t < KMAX; ++t)
            if (u32(t) < k) ids[u64(i) * k + t] = L.bid[t];
    }
}

// Compile-time dimension. The points come in a spatial order (xs, with
// perm[pos] the original id), and query lanes take consecutive positions, so
// a wave's queries are near one another.
//
// knn_seed: each query's k-th distance among the other points of its own
// SEED-position range (its neighbourhood in that order): an upper bound of
// its true k-th distance, +inf when the range holds fewer than k others.
constexpr u32 SEED = 4096;

template <int V, int DIM, int KMAX>
__global__ __launch_bounds__(GB) void knn_seed(const float* __restrict__ xs, const u32* __restrict__ perm, u32 n,
                                               MetricArgs ma, u32 k, float* __restrict__ kth0) {
    const u32 qp = blockIdx.x * GB + threadIdx.x;
    const u32 qc = qp < n ? qp : n - 1;
    float q[DIM];
#pragma unroll
    for (int t = 0; t < DIM; ++t) q[t] = xs[u64(qc) * DIM + t];
    // the wave's first lane decides both ranges, so candidates stay
    // wave-uniform: the wave's own 64 positions give a first bound (the k-th
    // distance within them, +inf if they hold fewer than k others), and the
    // SEED range is then scanned admitting only distances up to it. That
    // range contains the 64, so its k-th distance is at most the bound, and
    // the list's k-th is exactly the range's.
    const u32 first = __builtin_amdgcn_readfirstlane(qc);
    const u32 w0 = first / 64 * 64, h0 = first / SEED * SEED;
    KnnList<KMAX> L;
    L.init(k);
    knn_scan_static<V, DIM, KMAX>(xs, w0, min(n, w0 + 64), qc, q, ma, k, L);
    const float bound = L.kth;
    L.init(k);
    knn_scan_sorted<V, DIM, KMAX>(xs, perm, h0, min(n, h0 + SEED), qc, q, ma, k, bound, L);
    if (qp < n) kth0[qp] = L.kth;
}

// The scan proper: the candidates are split into gridDim.y chunks
// (blockIdx.y = chunk), so that a few thousand queries still fill the
// machine; each chunk admits only distances up to the query's seed bound
// (knn_scan_sorted). With one chunk the ids are written directly; otherwise
// each chunk's sorted (distance, id) list goes to pd / pi (row = position) and
// knn_merge combines them: the k best of the union are the k best of the
// per-chunk k best.
template <int V, int DIM, int KMAX>
__global__ __launch_bounds__(GB) void knn_scan_s(const float* __restrict__ xs, const u32* __restrict__ perm, u32 n,
                                                 MetricArgs ma, u32 k, const float* __restrict__ kth0,
                                                 u32* __restrict__ ids, float* __restrict__ pd,
                                                 u32* __restrict__ pi) {
    const u32 qp = blockIdx.x * GB + threadIdx.x;
    const u32 chunk = (n + gridDim.y - 1) / gridDim.y;
    const u32 jlo = blockIdx.y * chunk, jhi = min(n, jlo + chunk);
    const u32 qc = qp < n ? qp : n - 1;
    float q[DIM];
#pragma unroll
    for (int t = 0; t < DIM; ++t) q[t] = xs[u64(qc) * DIM + t];
    KnnList<KMAX> L;
    L.init(k);
    knn_scan_sorted<V, DIM, KMAX>(xs, perm, jlo, jhi, qc, q, ma, k, kth0[qc], L);
    if (qp < n) {
        if (gridDim.y == 1) {
            const u64 row = u64(perm[qp]) * k;
#pragma unroll
            for (int t = 0; t < KMAX; ++t)
                if (u32(t) < k) ids[row + t] = L.bid[t];
        } else {
            const u64 base = (u64(blockIdx.y) * n + qp) * k;
#pragma unroll
            for (int t = 0; t < KMAX; ++t)
                if (u32(t) < k) {
                    pd[base + t] = L.bd[t];
                    pi[base + t] = L.bid[t];
                }
        }
    }
}

// Merge of the per-chunk lists (each sorted by (distance, id); unused slots
// hold (+inf, ~0)), one lane per query position: the k smallest of the
// union, written to the query's original row perm[i].
__global__ __launch_bounds__(GB) void knn_merge(const float* __restrict__ pd, const u32* __restrict__ pi, u32 n,
                                                u32 k, u32 chunks, const u32* __restrict__ perm,
                                                u32* __restrict__ ids) {
    const u32 i = blockIdx.x * GB + threadIdx.x;
    if (i >= n) return;
    const u64 row = u64(perm[i]) * k;
    u32 cur[16];
    for (u32 c = 0; c < chunks; ++c) cur[c] = 0;
    for (u32 t = 0; t < k; ++t) {
        float bd = INFINITY;
        u32 bj = 0xffffffffu, bc = 0;
        for (u32 c = 0; c < chunks; ++c) {
            if (cur[c] >= k) continue;
            const u64 at = (u64(c) * n + i) * k + cur[c];
            const float d = pd[at];
            const u32 j = pi[at];
            if (d < bd || (d == bd && j < bj)) {
                bd = d;
                bj = j;
                bc = c;
            }
        }
        ids[row + t] = bj;
        ++cur[bc];
    }
}

// ---- kNN for k > 64: selection instead of a register list ---------------------
constexpr u32 KS = 256;   // threads per query

// 8-bit digit of the radix select at `shift` where the running rank `want`
// (1-based) falls, from a 256-bin histogram in LDS; updates want to the rank
// within that digit and returns the digit. Whole block.
__device__ __forceinline__ u32 select_digit(u32* hist, u32& want, u32* s_digit, u32* s_want) {
    __syncthreads();
    if (threadIdx.x < 64) {
        const u32 lane = threadIdx.x;
        u32 c[4], sum = 0;
#pragma unroll
        for (u32 t = 0; t < 4; ++t) {
            c[t] = hist[lane * 4 + t];
            sum += c[t];
        }
        u32 inc = sum;   // inclusive scan of the lane sums
#pragma unroll
        for (u32 off = 1; off < 64; off <<= 1) {
            const u32 o = __shfl_up(inc, off);
            if (lane >= off) inc += o;
        }
        u32 run = inc - sum;
#pragma unroll
        for (u32 t = 0; t < 4; ++t) {
            if (run < want && run + c[t] >= want) {
                *s_digit = lane * 4 + t;
                *s_want = want - run;
            }
            run += c[t];
        }
    }
    __syncthreads();
    want = *s_want;
    return *s_digit;
}

template <int V, int DIM>
__device__ __forceinline__ float knn_pair(const float* __restrict__ x, const float* q, const float* qp, u32 j,
                                          u32 dim, MetricArgs ma) {
    if constexpr (DIM > 0) {
        float c[DIM];
#pragma unroll
        for (int t = 0; t < DIM; ++t) c[t] = x[u64(j) * DIM + t];
        return pdist<V, DIM>(q, c, DIM, ma);
    } else {
        return pdist<V, 0>(qp, x + u64(j) * dim, dim, ma);
    }
}

// Per query (blockIdx.x): the k nearest j != q by (distance, id), written in
// any order as (distance bits << jb | j) keys with the query as payload.
template <int V, int DIM>
__global__ __launch_bounds__(KS) void knn_select(const float* __restrict__ x, u32 n, u32 dim_rt, MetricArgs ma,
                                                 u32 k, int jb, u64* __restrict__ keys, u32* __restrict__ owner) {
    constexpr u32 DS = DIM > 0 ? DIM : 1;
    __shared__ u32 hist[256];
    __shared__ u32 s_digit, s_want, s_count;
    const u32 dim = DIM > 0 ? u32(DIM) : dim_rt;
    const u32 qi = blockIdx.x;
    const float* qp = x + u64(qi) * dim;
    float q[DS];
    if constexpr (DIM > 0) {
#pragma unroll
        for (int t = 0; t < DIM; ++t) q[t] = qp[t];
    }
    // the k-th distance: four passes over the distance bits, most significant first
    u32 want = k, dpre = 0, dmask = 0;
    for (int shift = 24; shift >= 0; shift -= 8) {
        for (u32 t = threadIdx.x; t < 256; t += KS) hist[t] = 0;
        __syncthreads();
        for (u32 j = threadIdx.x; j < n; j += KS) {
            if (j == qi) continue;
            const u32 b = __float_as_uint(knn_pair<V, DIM>(x, q, qp, j, dim, ma));
            if ((b & dmask) == dpre) atomicAdd(&hist[(b >> shift) & 255u], 1u);
        }
        const u32 before = want;
        const u32 g = select_digit(hist, want, &s_digit, &s_want);
        (void)before;
        dpre |= g << shift;
        dmask |= 255u << shift;
    }
    // `want` of the candidates at distance D* = dpre are taken, the smallest ids
    // first (the reference's (distance, id) order); hist[g] of the last pass
    // holds how many there are
    const u32 ties = hist[dpre & 255u];
    u32 jmax = 0xffffffffu;
    if (want < ties) {
        u32 jpre = 0, jmask = 0;
        for (int shift = 24; shift >= 0; shift -= 8) {
            __syncthreads();
            for (u32 t = threadIdx.x; t < 256; t += KS) hist[t] = 0;
            __syncthreads();
            for (u32 j = threadIdx.x; j < n; j += KS) {
                if (j == qi || (j & jmask) != jpre) continue;
                if (__float_as_uint(knn_pair<V, DIM>(x, q, qp, j, dim, ma)) == dpre)
                    atomicAdd(&hist[(j >> shift) & 255u], 1u);
            }
            const u32 g = select_digit(hist, want, &s_digit, &s_want);
            jpre |= g << shift;
            jmask |= 255u << shift;
        }
        jmax = jpre;
    }
    if (threadIdx.x == 0) s_count = 0;
    __syncthreads();
    const u64 base = u64(qi) * k;
    for (u32 j = threadIdx.x; j < n; j += KS) {
        if (j == qi) continue;
        const u32 b = __float_as_uint(knn_pair<V, DIM>(x, q, qp, j, dim, ma));
        if (b < dpre || (b == dpre && j <= jmax)) {
            const u32 slot = atomicAdd(&s_count, 1u);
            if (slot < k) {
                keys[base + slot] = (u64(b) << jb) | j;
                owner[base + slot] = qi;
            }
        }
    }
}

template <int V, int DIM>
struct KnnSelectLaunch {
    static void run(u32 n, hipStream_t s, const float* x, u32 dim, MetricArgs ma, u32 k, int jb, u64* keys,
                    u32* owner) {
        knn_select<V, DIM><<<n, KS, 0, s>>>(x, n, dim, ma, k, jb, keys, owner);
    }
};

// (query << 32 | position) keys: the second sort groups the (distance, id)
// order by query
__global__ void knn_owner_keys(const u32* __restrict__ owner, u64 total, u64* __restrict__ keys) {
    for (u64 t = blockIdx.x * u64(blockDim.x) + threadIdx.x; t < total; t += u64(gridDim.x) * blockDim.x)
        keys[t] = (u64(owner[t]) << 32) | u32(t);
}

__global__ void knn_final_ids(const u64* __restrict__ by_query, const u64* __restrict__ by_dist, u64 total,
                              u64 jmask, u32* __restrict__ ids) {
    for (u64 t = blockIdx.x * u64(blockDim.x) + threadIdx.x; t < total; t += u64(gridDim.x) * blockDim.x)
        ids[t] = u32(by_dist[u32(by_query[t])] & jmask);
}

struct KnnChunks {
    u32 chunks = 1;          // candidate chunks (compile-time dimensions only)
    float* pd = nullptr;     // chunks x n x k partial lists
    u32* pi = nullptr;
    const float* xs = nullptr;   // the points in spatial order (compile-time dimensions)
    const u32* perm = nullptr;   // original id of each position
    float* kth0 = nullptr;       // per position: the seed bound
};

template <int V, int DIM, int KMAX>
void knn_one(u32 blocks, hipStream_t s, const float* x, u32 n, u32 dim, MetricArgs ma, u32 k, u32* ids,
             const KnnChunks& kc) {
    if constexpr (DIM > 0) {
        knn_seed<V, DIM, KMAX><<<blocks, GB, 0, s>>>(kc.xs, kc.perm, n, ma, k, kc.kth0);
        knn_scan_s<V, DIM, KMAX><<<dim3(blocks, kc.chunks), GB, 0, s>>>(kc.xs, kc.perm, n, ma, k, kc.kth0, ids,
                                                                        kc.pd, kc.pi);
        if (kc.chunks > 1) knn_merge<<<blocks, GB, 0, s>>>(kc.pd, kc.pi, n, k, kc.chunks, kc.perm, ids);
    } else {
        knn_scan<V, KMAX><<<blocks, GB, 0, s>>>(x, n, dim, ma, k, ids);
    }
}

template <int V, int DIM>
struct KnnLaunch {
    static void run(u32 kmax, u32 blocks, hipStream_t s, const float* x, u32 n, u32 dim,
                    MetricArgs ma, u32 k, u32* ids, const KnnChunks& kc) {
        switch (kmax) {
            case 8: knn_one<V, DIM, 8>(blocks, s, x, n, dim, ma, k, ids, kc); break;
            case 16: knn_one<V, DIM, 16>(blocks, s, x, n, dim, ma, k, ids, kc); break;
            case 32: knn_one<V, DIM, 32>(blocks, s, x, n, dim, ma, k, ids, kc); break;
            default: knn_one<V, DIM, 64>(blocks, s, x, n, dim, ma, k, ids, kc); break;
        }
    }
};

// ---- kNN graph ----------------------------------------------------------------
__global__ void emit_knn_keys(const u32* __restrict__ ids, u32 n, u32 k, int cb,
                              u64* __restrict__ keys) {
    const u64 total = u64(n) * k;
    for (u64 t = blockIdx.x * u64(blockDim.x) + threadIdx.x; t < total; t += u64(gridDim.x) * blockDim.x) {
        const u32 i = u32(t / k), j = ids[t];
        if (j == i || j >= n) {
            keys[2 * t] = keys[2 * t + 1] = ~0ull;   // sorts last, dropped
        } else {
            keys[2 * t] = (u64(i) << cb) | j;
            keys[2 * t + 1] = (u64(j) << cb) | i;
        }
    }
}

__global__ void mark_unique(const u64* __restrict__ keys, u64 m, u32* __restrict__ flag) {
    for (u64 t = blockIdx.x * u64(blockDim.x) + threadIdx.x; t < m; t += u64(gridDim.x) * blockDim.x)
        flag[t] = keys[t] != ~0ull && (t == 0 || keys[t] != keys[t - 1]);
}

template <int V, int DIM>
__global__ void knn_graph_fill(const float* __restrict__ x, u32 dim_rt, MetricArgs ma,
                               const u64* __restrict__ keys, const u32* __restrict__ flag,
                               const u32* __restrict__ pos, u64 m, int cb,
                               u32* __restrict__ col, float* __restrict__ w,
                               u32* __restrict__ deg) {
    const u32 dim = DIM > 0 ? u32(DIM) : dim_rt;
    const u64 mask = (u64(1) << cb) - 1;
    for (u64 t = blockIdx.x * u64(blockDim.x) + threadIdx.x; t < m; t += u64(gridDim.x) * blockDim.x) {
        if (!flag[t]) continue;
        const u32 r = u32(keys[t] >> cb), c = u32(keys[t] & mask);
        const u32 o = pos[t];
        col[o] = c;
        w[o] = pdist<V, DIM>(x + u64(r) * dim, x + u64(c) * dim, dim, ma);   // c(v, u), graph.hpp:301
        atomicAdd(deg + r, 1u);
    }
}

template <int V, int DIM>
struct KnnGraphFill {
    static void run(u32 blocks, hipStream_t s, const float* x, u32 dim, MetricArgs ma,
                    const u64* keys, const u32* flag, const u32* pos, u64 m, int cb, u32* col,
                    float* w, u32* deg) {
        knn_graph_fill<V, DIM><<<blocks, GB, 0, s>>>(x, dim, ma, keys, flag, pos, m, cb, col, w, deg);
    }
};

u32 blocks_for(u64 count, u32 per = GB) {
    return u32(std::max<u64>(1, std::min<u64>((count + per - 1) / per, 65536)));
}

void check_cloud(const float* pts, u32 n, u32 dim, int kind, float p, const char* what) {
    check_metric(kind, p);
    check_points(pts, u64(n) * dim, what);
    if (n && dim == 0) fail(TT_EINVAL, std::string(what) + ": dim must be >= 1");
}

}  // namespace


// CSR over n vertices from keys (row << cb | col) sorted ascending; deg[] holds
// the row counts. Writes row_ptr (n + 1) into ws and returns it.
static u64* row_ptr_from_deg(u32* deg, u32 n, Workspace& ws, hipStream_t s) {
    u64* rp = ws.get<u64>(u64(n) + 1);
    widen_u32<<<blocks_for(n), GB, 0, s>>>(deg, rp, n);
    TTG_LAUNCH_CHECK();
    exclusive_scan_u64(rp, rp, u64(n) + 1, nullptr, ws, s);
    return rp;
}

struct DevCsr {
    u64 nnz = 0;
    u64* row_ptr = nullptr;
    u32* col = nullptr;
    float* w = nullptr;
};

// ---- radius graph --------------------------------------------------------
DevCsr radius_graph_dev(const float* d_x, u32 n, u32 dim, int kind, float p, float tau,
                        Workspace& ws, hipStream_t s) {
    DevCsr g;
    u32* deg = ws.get<u32>(u64(n) + 1);
    fill_device_u32(deg, u64(n) + 1, 0u, s);
    if (n < 2) {
        g.row_ptr = row_ptr_from_deg(deg, n, ws, s);
        return g;
    }
    const MetricVariant v = variant_of(kind, p);
    const MetricArgs ma = metric_args(kind, p);
    const u64 nb = (u64(n) + GT - 1) / GT;
    const u64 tiles = nb * (nb + 1) / 2;
    const u32 blocks = u32(std::min<u64>(tiles, 16384));
    // the points in spatial order with a box per 64-point block
    float* xs = ws.get<float>(u64(n) * dim);
    u32* perm = ws.get<u32>(n);
    float* box = ws.get<float>(nb * 2 * dim);
    float* box16 = ws.get<float>((u64(n) + ROWS - 1) / ROWS * 2 * dim);   // per 16 positions
    {
        u32* dn = ws.get<u32>(1);
        fill_device_u32(dn, 1, n, s);
        spatial_order(d_x, n, dim, dn, ws.get<u64>(n), ws.get<u64>(n), ws.get<u64>(radix_scratch_entries(n)),
                      ws.get<u32>(spatial_ext_entries(dim)), xs, perm, s);
        spatial_boxes(xs, n, GT, dim, false, box, s);
        spatial_boxes(xs, n, ROWS, dim, false, box16, s);
    }
    unsigned long long* counter = ws.get<unsigned long long>(1);
    u64 cap = std::min<u64>(u64(n) * (n - 1) / 2, std::max<u64>(u64(1) << 20, u64(n) * 32));
    u64 m = 0;
    for (int attempt = 0; attempt < 2; ++attempt) {
        u32* ei = ws.get<u32>(cap);
        u32* ej = ws.get<u32>(cap);
        float* ew = ws.get<float>(cap);
        fill_device_u32(reinterpret_cast<u32*>(counter), 2, 0u, s);
        dispatch_metric_dim<RadiusLaunch>(v, dim, blocks, s, xs, perm, box, box16, n, dim, ma, tau, tiles, cap,
                                          counter, ei, ej, ew);
        TTG_LAUNCH_CHECK();
        TTG_HIP(hipMemcpyAsync(&m, counter, sizeof(u64), hipMemcpyDeviceToHost, s));
        TTG_HIP(hipStreamSynchronize(s));
        if (m <= cap) {
            const int cb = std::max(1, bit_length(n - 1));
            const u64 nnz = 2 * m;
            u64* keys = ws.get<u64>(nnz);
            u64* kalt = ws.get<u64>(nnz);
            u32* vals = ws.get<u32>(nnz);
            u32* valt = ws.get<u32>(nnz);
            if (m) {
                emit_both<<<blocks_for(m), GB, 0, s>>>(ei, ej, ew, m, cb, keys, vals, deg);
                TTG_LAUNCH_CHECK();
                radix_sort_u64(keys, vals, nnz, 0, 2 * cb, kalt, valt, ws, s);
            }
            g.nnz = nnz;
            g.col = ws.get<u32>(nnz);
            g.w = ws.get<float>(nnz);
            if (nnz) {
                split_keys<<<blocks_for(nnz), GB, 0, s>>>(keys, vals, nnz, (u64(1) << cb) - 1, g.col, g.w);
                TTG_LAUNCH_CHECK();
            }
            g.row_ptr = row_ptr_from_deg(deg, n, ws, s);
            return g;
        }
        cap = m;   // exact size known now
    }
    fail(TT_EDEVICE, "radius_graph: edge count changed between passes");
}

// ---- kNN --------------------------------------------------------------------
// k > 64: per-query selection, then the lists ordered by (query, distance, id)
void knn_large_k(const float* d_x, u32 n, u32 dim, int kind, float p, u32 kk, u32* d_ids, Workspace& ws,
                 hipStream_t s) {
    const u64 total = u64(n) * kk;
    if (total >= (u64(1) << 32)) fail(TT_ELENGTH, "knn: n * k must be below 2^32");
    const int jb = std::max(1, bit_length(n - 1));
    u64* keys = ws.get<u64>(total);
    u64* kalt = ws.get<u64>(total);
    u32* owner = ws.get<u32>(total);
    u32* oalt = ws.get<u32>(total);
    const MetricVariant v = variant_of(kind, p);
    const MetricArgs ma = metric_args(kind, p);
    dispatch_metric_dim<KnnSelectLaunch>(v, dim, n, s, d_x, dim, ma, kk, jb, keys, owner);
    TTG_LAUNCH_CHECK();
    // (distance, id) ascending, the query carried along; then stably by query
    radix_sort_u64(keys, owner, total, 0, 31 + jb, kalt, oalt, ws, s);
    u64* qkeys = ws.get<u64>(total);
    knn_owner_keys<<<blocks_for(total), GB, 0, s>>>(owner, total, qkeys);
    TTG_LAUNCH_CHECK();
    radix_sort_u64(qkeys, nullptr, total, 32, 32 + std::max(1, bit_length(n - 1)), kalt, nullptr, ws, s);
    knn_final_ids<<<blocks_for(total), GB, 0, s>>>(qkeys, keys, total, (u64(1) << jb) - 1, d_ids);
    TTG_LAUNCH_CHECK();
}

u32 knn_dev(const float* d_x, u32 n, u32 dim, int kind, float p, u32 k, u32* d_ids, Workspace& ws,
            hipStream_t s) {
    const u32 kk = n ? std::min<u32>(k, n - 1) : 0;
    if (!kk) return 0;
    if (kk > 64) {
        knn_large_k(d_x, n, dim, kind, p, kk, d_ids, ws, s);
        return kk;
    }
    const u32 kmax = kk <= 8 ? 8 : kk <= 16 ? 16 : kk <= 32 ? 32 : 64;
    const MetricVariant v = variant_of(kind, p);
    const MetricArgs ma = metric_args(kind, p);
    const u32 blocks = (n + GB - 1) / GB;
    // enough candidate chunks for about 8192 waves (a full machine), at most
    // 16, each chunk at least 4096 candidates. More chunks measured slower on
    // n = 2e5 (each chunk refills its list from empty, and a wave pays for an
    // insertion whenever one of its lanes makes one): 46.5 ms at 8192 waves,
    // 50.6 at 16384, 54.7 with one chunk.
    KnnChunks kc;
    const u32 waves = (n + 63) / 64;
    kc.chunks = std::max<u32>(1, std::min<u32>({16u, (8192 + waves - 1) / waves, std::max<u32>(1, n / 4096)}));
    if (kc.chunks > 1) {
        kc.pd = ws.get<float>(u64(kc.chunks) * n * kk);
        kc.pi = ws.get<u32>(u64(kc.chunks) * n * kk);
    }
    if (dim == 1 || dim == 2 || dim == 3 || dim == 4 || dim == 8 || dim == 16) {   // the compile-time scans
        float* xs = ws.get<float>(u64(n) * dim);
        u32* perm = ws.get<u32>(n);
        u32* dn = ws.get<u32>(1);
        fill_device_u32(dn, 1, n, s);
        spatial_order(d_x, n, dim, dn, ws.get<u64>(n), ws.get<u64>(n), ws.get<u64>(radix_scratch_entries(n)),
                      ws.get<u32>(spatial_ext_entries(dim)), xs, perm, s);
        kc.xs = xs;
        kc.perm = perm;
        kc.kth0 = ws.get<float>(n);
    }
    dispatch_metric_dim<KnnLaunch>(v, dim, kmax, blocks, s, d_x, n, dim, ma, kk, d_ids, kc);
    TTG_LAUNCH_CHECK();
    return kk;
}

// ---- kNN graph --------------------------------------------------------------
DevCsr knn_graph_dev(const float* d_x, u32 n, u32 dim, int kind, float p, u32 k, const u32* d_ids,
                     Workspace& ws, hipStream_t s) {
    DevCsr g;
    u32* deg = ws.get<u32>(u64(n) + 1);
    fill_device_u32(deg, u64(n) + 1, 0u, s);
    const u64 m = 2 * u64(n) * k;
    if (!m) {
        g.row_ptr = row_ptr_from_deg(deg, n, ws, s);
        return g;
    }
    const int cb = std::max(1, bit_length(n - 1));
    u64* keys = ws.get<u64>(m);
    u64* kalt = ws.get<u64>(m);
    emit_knn_keys<<<blocks_for(u64(n) * k), GB, 0, s>>>(d_ids, n, k, cb, keys);
    TTG_LAUNCH_CHECK();
    radix_sort_u64(keys, nullptr, m, 0, std::min(64, 2 * cb + 1), kalt, nullptr, ws, s);
    u32* flag = ws.get<u32>(m);
    u32* pos = ws.get<u32>(m);
    mark_unique<<<blocks_for(m), GB, 0, s>>>(keys, m, flag);
    TTG_LAUNCH_CHECK();
    u32* total = ws.get<u32>(1);
    exclusive_scan_u32(flag, pos, m, total, ws, s);
    u32 nnz = 0;
    TTG_HIP(hipMemcpyAsync(&nnz, total, sizeof(u32), hipMemcpyDeviceToHost, s));
    TTG_HIP(hipStreamSynchronize(s));
    g.nnz = nnz;
    g.col = ws.get<u32>(nnz);
    g.w = ws.get<float>(nnz);
    const MetricVariant v = variant_of(kind, p);
    const MetricArgs ma = metric_args(kind, p);
    dispatch_metric_dim<KnnGraphFill>(v, dim, blocks_for(m), s, d_x, dim, ma, keys, flag, pos, m, cb,
                                      g.col, g.w, deg);
    TTG_LAUNCH_CHECK();
    g.row_ptr = row_ptr_from_deg(deg, n, ws, s);
    return g;
}

}  // namespace ttg

using namespace ttg;

namespace {

// Copies a device CSR to library-malloc'd host arrays (released by tt_free).
void csr_to_host(const DevCsr& g, u32 n, uint64_t* row_ptr, uint64_t* nnz, uint32_t** col,
                 float** w, hipStream_t s) {
    *nnz = g.nnz;
    TTG_HIP(hipMemcpyAsync(row_ptr, g.row_ptr, (u64(n) + 1) * sizeof(u64), hipMemcpyDeviceToHost, s));
    *col = static_cast<uint32_t*>(std::malloc(std::max<u64>(1, g.nnz) * sizeof(u32)));
    *w = static_cast<float*>(std::malloc(std::max<u64>(1, g.nnz) * sizeof(float)));
    if (!*col || !*w) fail(TT_EDEVICE, "out of host memory");
    if (g.nnz) {
        TTG_HIP(hipMemcpyAsync(*col, g.col, g.nnz * sizeof(u32), hipMemcpyDeviceToHost, s));
        TTG_HIP(hipMemcpyAsync(*w, g.w, g.nnz * sizeof(float), hipMemcpyDeviceToHost, s));
    }
    TTG_HIP(hipStreamSynchronize(s));
}

}  // namespace

extern "C" {

tt_status tt_radius_graph(const float* pts, uint32_t n, uint32_t dim, int kind, float p, float tau,
                          uint64_t* row_ptr, uint64_t* nnz, uint32_t** col, float** w) {
    return guarded([&] {
        check_cloud(pts, n, dim, kind, p, "tt_radius_graph");
        if (!row_ptr || !nnz || !col || !w) fail(TT_EINVAL, "tt_radius_graph: null output");
        if (std::isnan(tau) || tau < 0.0f)
            fail(TT_EINVAL, "persistence: threshold must be non-negative or +inf");
        hipStream_t s = nullptr;
        Workspace ws;
        float* dx = ws.get<float>(u64(n) * dim);
        if (u64(n) * dim)
            TTG_HIP(hipMemcpyAsync(dx, pts, u64(n) * dim * sizeof(float), hipMemcpyHostToDevice, s));
        const DevCsr g = radius_graph_dev(dx, n, dim, kind, p, tau, ws, s);
        csr_to_host(g, n, row_ptr, nnz, col, w, s);
    });
}

tt_status tt_knn_exact(const float* pts, uint32_t n, uint32_t dim, int kind, float p, uint32_t k,
                       uint32_t* k_out, uint32_t* ids) {
    return guarded([&] {
        check_cloud(pts, n, dim, kind, p, "tt_knn_exact");
        if (!k_out) fail(TT_EINVAL, "tt_knn_exact: null k_out");
        const u32 kk = n ? std::min<u32>(k, n - 1) : 0;
        *k_out = kk;
        if (!kk) return;
        if (!ids) fail(TT_EINVAL, "tt_knn_exact: null ids");
        hipStream_t s = nullptr;
        Workspace ws;
        float* dx = ws.get<float>(u64(n) * dim);
        u32* di = ws.get<u32>(u64(n) * kk);
        TTG_HIP(hipMemcpyAsync(dx, pts, u64(n) * dim * sizeof(float), hipMemcpyHostToDevice, s));
        knn_dev(dx, n, dim, kind, p, k, di, ws, s);
        TTG_HIP(hipMemcpyAsync(ids, di, u64(n) * kk * sizeof(u32), hipMemcpyDeviceToHost, s));
        TTG_HIP(hipStreamSynchronize(s));
    });
}

tt_status tt_knn_graph(const float* pts, uint32_t n, uint32_t dim, int kind, float p, uint32_t k,
                       const uint32_t* ids, uint64_t* row_ptr, uint64_t* nnz, uint32_t** col,
                       float** w) {
    return guarded([&] {
        check_cloud(pts, n, dim, kind, p, "tt_knn_graph");
        if (!row_ptr || !nnz || !col || !w) fail(TT_EINVAL, "tt_knn_graph: null output");
        if (u64(n) * k && !ids) fail(TT_EINVAL, "tt_knn_graph: null ids");
        hipStream_t s = nullptr;
        Workspace ws;
        float* dx = ws.get<float>(u64(n) * dim);
        u32* di = ws.get<u32>(u64(n) * k);
        if (u64(n) * dim)
            TTG_HIP(hipMemcpyAsync(dx, pts, u64(n) * dim * sizeof(float), hipMemcpyHostToDevice, s));
        if (u64(n) * k)
            TTG_HIP(hipMemcpyAsync(di, ids, u64(n) * k * sizeof(u32), hipMemcpyHostToDevice, s));
        const DevCsr g = knn_graph_dev(dx, n, dim, kind, p, k, di, ws, s);
        csr_to_host(g, n, row_ptr, nnz, col, w, s);
    });
}

// knn_lists_exact + knn_graph in one device pass (the persistence() path,
// turbotops.hpp:275-277): the lists never leave HBM.
tt_status tt_knn_graph_exact(const float* pts, uint32_t n, uint32_t dim, int kind, float p,
                             uint32_t k, uint64_t* row_ptr, uint64_t* nnz, uint32_t** col,
                             float** w) {
    return guarded([&] {
        check_cloud(pts, n, dim, kind, p, "tt_knn_graph_exact");
        if (k == 0) fail(TT_EINVAL, "persistence: k must be >= 1");
        if (!row_ptr || !nnz || !col || !w) fail(TT_EINVAL, "tt_knn_graph_exact: null output");
        hipStream_t s = nullptr;
        Workspace ws;
        const u32 kk = n ? std::min<u32>(k, n - 1) : 0;
        float* dx = ws.get<float>(u64(n) * dim);
        u32* di = ws.get<u32>(u64(n) * kk);
        if (u64(n) * dim)
            TTG_HIP(hipMemcpyAsync(dx, pts, u64(n) * dim * sizeof(float), hipMemcpyHostToDevice, s));
        knn_dev(dx, n, dim, kind, p, k, di, ws, s);
        const DevCsr g = knn_graph_dev(dx, n, dim, kind, p, kk, di, ws, s);
        csr_to_host(g, n, row_ptr, nnz, col, w, s);
    });
}

}  // extern "C"
