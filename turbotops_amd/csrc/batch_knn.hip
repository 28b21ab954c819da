// SPDX-License-Identifier: MIT
// batch_knn.hip — the kNN stage of the batched front end: for many point
// clouds of uneven size and dimension, each cloud's exact k-nearest-neighbour
// lists (knn_lists_exact, graph.hpp:211-225) and their symmetric union as a
// CSR graph with recomputed weights (knn_graph, graph.hpp:286-304;
// CsrGraph::from_rows, core.hpp:113-134), in one launch sequence recorded
// once into a HIP graph. Each cloud's graph is what persistence(cloud,
// graph="knn", k) reduces.
//
//   spatial     per cloud: coordinate extents (bknn_extent), a grid code per
//   order       point (bknn_codes), one stable sort of (cloud, code, index)
//               keys over every cloud, the points copied in that order
//               (bknn_gather) with each position's original index
//   bknn_scan   one query per lane, one cloud per workgroup of 256 queries;
//               one launch per dimension group (2, 3, 8, other), so each
//               runs the list code specialised for its dimension (knn.hpp).
//               Compile-time dimensions follow the cloud's spatial order: a
//               wave's 64 neighbouring positions bound each query's k-th
//               distance, and the cloud is scanned admitting only distances up
//               to it (every pair still measured; ids are the original ones)
//   bknn_keys   both directions of every list entry as (cloud, row, col) keys
//   radix sort  one sort over every cloud (cloud-major, rows ascending, then
//               neighbour ids ascending: from_rows' order)
//   bknn_mark   first copy of each key; a scan gives its CSR slot
//   bknn_fill   neighbour id and weight c(row, col) per slot, row degrees
//   scan        row offsets over every cloud's rows at once
// Fixed grids and counts known at creation (every cloud's list length is
// min(k, n_c - 1)), so the sequence replays from a graph; the host then
// copies the offsets, ids and weights once. Clouds whose list is longer than
// 64 (k > 64 and n_c > 65) take the single-cloud selection path
// (knn_large_k) in place of the register-list scan, one cloud after another;
// its temporaries are allocated per run, so a batch holding such a cloud runs
// its sequence directly instead of from the recorded graph.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "frontend.hpp"
#include "knn.hpp"

namespace ttg {

namespace {

constexpr u32 QB = 256;   // queries per workgroup

struct KnnBatchShape {
    const u32* n;
    const u32* dim;
    const u64* pt_off;     // nclouds + 1
    const u64* row_base;   // nclouds + 1
    const u64* list_off;   // nclouds + 1: sum of n_c * k_c before cloud c
    const u32* row_cloud;  // total_rows
    u32 k;
    u64 total_rows;
};

__device__ __forceinline__ u32 list_len(u32 k, u32 n) { return n >= 2 ? min(k, n - 1) : 0u; }

template <int V, int DIM, int KMAX>
__global__ __launch_bounds__(QB) void bknn_scan(KnnBatchShape b, MetricArgs ma, const float* __restrict__ pts,
                                                const float* __restrict__ xs, const u32* __restrict__ perm,
                                                const u32* __restrict__ qb_cloud, const u32* __restrict__ qb_first,
                                                u64 qb0, u32* __restrict__ ids) {
    const u64 blk = qb0 + blockIdx.x;
    const u32 c = __builtin_amdgcn_readfirstlane(qb_cloud[blk]);
    const u32 n = b.n[c];
    const u32 kc = list_len(b.k, n);   // >= 1: only clouds with n >= 2 have query blocks
    const u32 i = qb_first[blk] + threadIdx.x;
    const u32 qi = i < n ? i : n - 1;
    KnnList<KMAX> L;
    L.init(kc);
    if constexpr (DIM > 0) {
        // the cloud in its spatial order (xs; perm = original index per
        // position): lane = position, so a wave's queries are neighbours. The
        // wave's own 64 positions give a first bound of each query's k-th
        // distance (+inf if they hold fewer than k others); the whole cloud is
        // then scanned admitting only distances up to it (knn_scan_sorted), the
        // list's ids and ties being the original ids.
        const float* __restrict__ x = xs + b.pt_off[c];
        const u32* __restrict__ pc = perm + b.row_base[c];
        float q[DIM];
#pragma unroll
        for (int t = 0; t < DIM; ++t) q[t] = x[u64(qi) * DIM + t];
        const u32 w0 = __builtin_amdgcn_readfirstlane(qi) / 64 * 64;
        knn_scan_static<V, DIM, KMAX>(x, w0, min(n, w0 + 64), qi, q, ma, kc, L);
        const float bound = L.kth;
        L.init(kc);
        knn_scan_sorted<V, DIM, KMAX>(x, pc, 0, n, qi, q, ma, kc, bound, L);
        if (i < n) {
            u32* out = ids + b.list_off[c] + u64(pc[i]) * kc;
#pragma unroll
            for (int t = 0; t < KMAX; ++t)
                if (u32(t) < kc) out[t] = L.bid[t];
        }
    } else {
        const float* __restrict__ x = pts + b.pt_off[c];
        knn_scan_runtime<V, KMAX>(x, b.dim[c], 0, n, qi, ma, kc, L);
        if (i < n) {
            u32* out = ids + b.list_off[c] + u64(i) * kc;
#pragma unroll
            for (int t = 0; t < KMAX; ++t)
                if (u32(t) < kc) out[t] = L.bid[t];
        }
    }
}

// ---- each cloud's spatial order (compile-time dimension groups) -------------
// Order-preserving map of float bits to unsigned (finite values).
__device__ __forceinline__ u32 bk_ord(float f) {
    const u32 u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float bk_unord(u32 u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

constexpr u32 EXT_STRIDE = 64;   // ext entries per cloud: 2 per coordinate, at most 32 coordinates

// One workgroup per cloud: the order bits of each coordinate's smallest and
// largest finite value over the cloud's points.
__global__ __launch_bounds__(256) void bknn_extent(KnnBatchShape b, const float* __restrict__ pts,
                                                   u32* __restrict__ ext) {
    __shared__ u32 slo[4], shi[4];
    const u32 c = blockIdx.x, n = b.n[c], dim = b.dim[c];
    const u32 kdim = min(dim, 32u);
    const float* __restrict__ x = pts + b.pt_off[c];
    const u32 lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (u32 k = 0; k < kdim; ++k) {
        u32 lo = 0xffffffffu, hi = 0u;
        for (u32 v = threadIdx.x; v < n; v += 256) {
            const float f = x[u64(v) * dim + k];
            if (__builtin_isfinite(f)) {
                lo = min(lo, bk_ord(f));
                hi = max(hi, bk_ord(f));
            }
        }
        for (int off = 32; off > 0; off >>= 1) {
            lo = min(lo, u32(__shfl_xor(int(lo), off)));
            hi = max(hi, u32(__shfl_xor(int(hi), off)));
        }
        if (lane == 0) {
            slo[w] = lo;
            shi[w] = hi;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (u32 q = 1; q < 4; ++q) {
                lo = min(lo, slo[q]);
                hi = max(hi, shi[q]);
            }
            ext[u64(c) * EXT_STRIDE + 2 * k] = lo;
            ext[u64(c) * EXT_STRIDE + 2 * k + 1] = hi;
        }
        __syncthreads();
    }
}

// key of point r: (cloud << (32 + codebits)) | (grid code << 32) | local index,
// the code interleaving `bits` bits of the cell of each of the first kuse
// coordinates over the cloud's extent.
__global__ void bknn_codes(KnnBatchShape b, const float* __restrict__ pts, const u32* __restrict__ ext,
                           u32 codebits, u64* __restrict__ keys) {
    for (u64 r = blockIdx.x * u64(blockDim.x) + threadIdx.x; r < b.total_rows; r += u64(gridDim.x) * blockDim.x) {
        const u32 c = b.row_cloud[r];
        const u32 v = u32(r - b.row_base[c]);
        const u32 dim = b.dim[c];
        const u32 kuse = min(min(dim, 32u), codebits);
        const u32 bits = max(1u, min(10u, codebits / kuse));
        const float cells = float(1u << bits);
        const float* __restrict__ p = pts + b.pt_off[c] + u64(v) * dim;
        u32 cell[32];
        for (u32 k = 0; k < kuse; ++k) {
            const float lo = bk_unord(ext[u64(c) * EXT_STRIDE + 2 * k]);
            const float hi = bk_unord(ext[u64(c) * EXT_STRIDE + 2 * k + 1]);
            const float f = p[k];
            u32 cl = 0;
            if (hi > lo && f > lo) {   // false for NaN and for an empty extent
                const float t = (f - lo) / (hi - lo) * cells;
                cl = t >= cells ? (1u << bits) - 1u : u32(t);
            }
            cell[k] = cl;
        }
        u64 code = 0;
        for (int bt = int(bits) - 1; bt >= 0; --bt)
            for (u32 k = 0; k < kuse; ++k) code = (code << 1) | ((cell[k] >> bt) & 1u);
        keys[r] = (u64(c) << (32 + codebits)) | (code << 32) | v;
    }
}

// Sorted position r (global; the cloud's positions stay within its own rows,
// the sort being stable with the cloud in the top bits): the point's copy in
// xs and its local original index.
__global__ void bknn_gather(KnnBatchShape b, const float* __restrict__ pts, const u64* __restrict__ sorted,
                            float* __restrict__ xs, u32* __restrict__ perm) {
    for (u64 r = blockIdx.x * u64(blockDim.x) + threadIdx.x; r < b.total_rows; r += u64(gridDim.x) * blockDim.x) {
        const u32 c = b.row_cloud[r];
        const u32 v = u32(sorted[r]);
        const u32 dim = b.dim[c];
        const u64 lp = r - b.row_base[c];
        for (u32 k = 0; k < dim; ++k) xs[b.pt_off[c] + lp * dim + k] = pts[b.pt_off[c] + u64(v) * dim + k];
        perm[r] = v;
    }
}

// Both directions of every list entry: (cloud << 2rb | row << rb | col).
__global__ void bknn_keys(KnnBatchShape b, const u32* __restrict__ ids, int rb, u64* __restrict__ keys) {
    for (u64 r = blockIdx.x * u64(blockDim.x) + threadIdx.x; r < b.total_rows; r += u64(gridDim.x) * blockDim.x) {
        const u32 c = b.row_cloud[r];
        const u32 kc = list_len(b.k, b.n[c]);
        const u64 i = r - b.row_base[c];
        const u64 cloud_bits = u64(c) << (2 * rb);
        const u64 base = b.list_off[c] + i * kc;
        for (u32 t = 0; t < kc; ++t) {
            const u64 j = ids[base + t];
            keys[2 * (base + t)] = cloud_bits | (i << rb) | j;
            keys[2 * (base + t) + 1] = cloud_bits | (j << rb) | i;
        }
    }
}

__global__ void bknn_mark(const u64* __restrict__ keys, u64 m, u32* __restrict__ flag) {
    for (u64 t = blockIdx.x * u64(blockDim.x) + threadIdx.x; t < m; t += u64(gridDim.x) * blockDim.x)
        flag[t] = t == 0 || keys[t] != keys[t - 1];
}

// The weight is recomputed as c(row, col) (graph.hpp:301), with the cloud's
// runtime dimension: pdist<V, 0> sums in the same order as the compile-time
// forms, so the weights are the single-cloud graph's bit for bit.
template <int V>
__global__ void bknn_fill(KnnBatchShape b, MetricArgs ma, const float* __restrict__ pts,
                          const u64* __restrict__ keys, const u32* __restrict__ flag, const u32* __restrict__ pos,
                          u64 m, int rb, u32* __restrict__ col, float* __restrict__ w, u32* __restrict__ deg) {
    const u64 mask = (u64(1) << rb) - 1;
    for (u64 t = blockIdx.x * u64(blockDim.x) + threadIdx.x; t < m; t += u64(gridDim.x) * blockDim.x) {
        if (!flag[t]) continue;
        const u64 key = keys[t];
        const u32 c = u32(key >> (2 * rb));
        const u32 r = u32((key >> rb) & mask), j = u32(key & mask);
        const u32 dim = b.dim[c];
        const float* __restrict__ x = pts + b.pt_off[c];
        const u32 o = pos[t];
        col[o] = j;
        w[o] = pdist<V, 0>(x + u64(r) * dim, x + u64(j) * dim, dim, ma);
        atomicAdd(deg + b.row_base[c] + r, 1u);
    }
}

constexpr int kDimGroups = 4;
constexpr u32 kGroupDim[kDimGroups] = {2, 3, 8, 0};
inline int dim_group(u32 d) { return d == 2 ? 0 : d == 3 ? 1 : d == 8 ? 2 : 3; }

template <int V, int KMAX>
void launch_scan_groups(const u64* gq, hipStream_t s, const KnnBatchShape& b, MetricArgs ma, const float* pts,
                        const float* xs, const u32* perm, const u32* qb_cloud, const u32* qb_first, u32* ids) {
    for (int g = 0; g < kDimGroups; ++g) {
        const u64 q0 = gq[g], q1 = gq[g + 1];
        if (q1 <= q0) continue;
        const u32 grid = u32(q1 - q0);
        switch (kGroupDim[g]) {
            case 2: bknn_scan<V, 2, KMAX><<<grid, QB, 0, s>>>(b, ma, pts, xs, perm, qb_cloud, qb_first, q0, ids); break;
            case 3: bknn_scan<V, 3, KMAX><<<grid, QB, 0, s>>>(b, ma, pts, xs, perm, qb_cloud, qb_first, q0, ids); break;
            case 8: bknn_scan<V, 8, KMAX><<<grid, QB, 0, s>>>(b, ma, pts, xs, perm, qb_cloud, qb_first, q0, ids); break;
            default: bknn_scan<V, 0, KMAX><<<grid, QB, 0, s>>>(b, ma, pts, xs, perm, qb_cloud, qb_first, q0, ids); break;
        }
    }
}

template <int V>
void launch_scan_k(u32 kmax, const u64* gq, hipStream_t s, const KnnBatchShape& b, MetricArgs ma, const float* pts,
                   const float* xs, const u32* perm, const u32* qb_cloud, const u32* qb_first, u32* ids) {
    switch (kmax) {
        case 8: launch_scan_groups<V, 8>(gq, s, b, ma, pts, xs, perm, qb_cloud, qb_first, ids); break;
        case 16: launch_scan_groups<V, 16>(gq, s, b, ma, pts, xs, perm, qb_cloud, qb_first, ids); break;
        case 32: launch_scan_groups<V, 32>(gq, s, b, ma, pts, xs, perm, qb_cloud, qb_first, ids); break;
        default: launch_scan_groups<V, 64>(gq, s, b, ma, pts, xs, perm, qb_cloud, qb_first, ids); break;
    }
}

void launch_scan(MetricVariant v, u32 kmax, const u64* gq, hipStream_t s, const KnnBatchShape& b, MetricArgs ma,
                 const float* pts, const float* xs, const u32* perm, const u32* qb_cloud, const u32* qb_first,
                 u32* ids) {
    switch (v) {
        case MV_L1: launch_scan_k<MV_L1>(kmax, gq, s, b, ma, pts, xs, perm, qb_cloud, qb_first, ids); break;
        case MV_L2: launch_scan_k<MV_L2>(kmax, gq, s, b, ma, pts, xs, perm, qb_cloud, qb_first, ids); break;
        case MV_LINF: launch_scan_k<MV_LINF>(kmax, gq, s, b, ma, pts, xs, perm, qb_cloud, qb_first, ids); break;
        case MV_LP_INT: launch_scan_k<MV_LP_INT>(kmax, gq, s, b, ma, pts, xs, perm, qb_cloud, qb_first, ids); break;
        default: launch_scan_k<MV_LP_POW>(kmax, gq, s, b, ma, pts, xs, perm, qb_cloud, qb_first, ids); break;
    }
}

void launch_fill(MetricVariant v, u32 grid, hipStream_t s, const KnnBatchShape& b, MetricArgs ma, const float* pts,
                 const u64* keys, const u32* flag, const u32* pos, u64 m, int rb, u32* col, float* w, u32* deg) {
    switch (v) {
        case MV_L1: bknn_fill<MV_L1><<<grid, 256, 0, s>>>(b, ma, pts, keys, flag, pos, m, rb, col, w, deg); break;
        case MV_L2: bknn_fill<MV_L2><<<grid, 256, 0, s>>>(b, ma, pts, keys, flag, pos, m, rb, col, w, deg); break;
        case MV_LINF: bknn_fill<MV_LINF><<<grid, 256, 0, s>>>(b, ma, pts, keys, flag, pos, m, rb, col, w, deg); break;
        case MV_LP_INT:
            bknn_fill<MV_LP_INT><<<grid, 256, 0, s>>>(b, ma, pts, keys, flag, pos, m, rb, col, w, deg);
            break;
        default: bknn_fill<MV_LP_POW><<<grid, 256, 0, s>>>(b, ma, pts, keys, flag, pos, m, rb, col, w, deg); break;
    }
}

u32 grid_for(u64 count, u32 per = 256) {
    return u32(std::max<u64>(1, std::min<u64>((count + per - 1) / per, 65536)));
}

}  // namespace

struct BatchKnnState {
    u32 nclouds = 0, k = 0, kmax = 8;
    int kind = 1;
    float p = 2.0f;
    std::vector<u32> large;   // clouds with a list longer than 64: the selection path
    Workspace ws_large;       // its temporaries, for one run
    int rb = 1, cb = 0;
    MetricVariant variant = MV_L2;
    MetricArgs ma{};
    std::vector<u32> n, dim;
    std::vector<u64> pt_off, row_base, list_off, gq;
    u64 total_pts = 0, total_rows = 0, total_lists = 0, m = 0;
    Workspace ws;
    float *d_pts = nullptr, *d_w = nullptr;
    u32 *d_ids = nullptr, *d_flag = nullptr, *d_pos = nullptr, *d_col = nullptr, *d_deg = nullptr,
        *d_total = nullptr, *d_count = nullptr, *d_stmp = nullptr, *d_n = nullptr,
        *d_dim = nullptr, *d_row_cloud = nullptr, *d_qb_cloud = nullptr, *d_qb_first = nullptr,
        *d_rp = nullptr;   // row offsets over every cloud's rows (total_rows + 1)
    u64 *d_keys = nullptr, *d_alt = nullptr, *d_pt_off = nullptr, *d_row_base = nullptr, *d_list_off = nullptr;
    u64* d_sort = nullptr;   // radix sort scratch (radix_scratch_entries)
    // each cloud's spatial order: keys and their sort, the sorted copy of the
    // points, the original index per position, extents, the point count
    u64 *d_skeys = nullptr, *d_salt = nullptr, *d_ssort = nullptr;
    float* d_xs = nullptr;
    u32 *d_perm = nullptr, *d_ext = nullptr, *d_nrows = nullptr;
    u32 codebits = 23;
    KnnBatchShape shape{};
    float* h_pts = nullptr;
    std::vector<u32> h_rp, h_col, lists;
    std::vector<float> h_w;
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    float last_ms = 0.0f;
    bool ran = false;

    ~BatchKnnState() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (t0) (void)hipEventDestroy(t0);
        if (t1) (void)hipEventDestroy(t1);
        if (stream) (void)hipStreamDestroy(stream);
        if (h_pts) (void)hipHostFree(h_pts);
    }

    void enqueue() {
        hipStream_t s = stream;
        if (!m) {   // no cloud has two points: every row is empty
            fill_device_u32(d_rp, total_rows + 1, 0u, s);
            return;
        }
        bknn_extent<<<nclouds, 256, 0, s>>>(shape, d_pts, d_ext);
        bknn_codes<<<grid_for(total_rows), 256, 0, s>>>(shape, d_pts, d_ext, codebits, d_skeys);
        TTG_LAUNCH_CHECK();
        const u64* order = radix_sort_u64_counted(d_skeys, std::max<u64>(total_rows, 1), d_nrows, 32, 64, d_salt,
                                                  d_ssort, s);
        bknn_gather<<<grid_for(total_rows), 256, 0, s>>>(shape, d_pts, order, d_xs, d_perm);
        TTG_LAUNCH_CHECK();
        launch_scan(variant, kmax, gq.data(), s, shape, ma, d_pts, d_xs, d_perm, d_qb_cloud, d_qb_first, d_ids);
        TTG_LAUNCH_CHECK();
        for (u32 c : large)
            knn_large_k(d_pts + pt_off[c], n[c], dim[c], kind, p, std::min(k, n[c] - 1), d_ids + list_off[c],
                        ws_large, s);
        bknn_keys<<<grid_for(total_rows), 256, 0, s>>>(shape, d_ids, rb, d_keys);
        TTG_LAUNCH_CHECK();
        const u64* sorted = radix_sort_u64_counted(d_keys, m, d_count, 0, cb + 2 * rb, d_alt, d_sort, s);
        bknn_mark<<<grid_for(m), 256, 0, s>>>(sorted, m, d_flag);
        TTG_LAUNCH_CHECK();
        exclusive_scan_u32_buf(d_flag, d_pos, m, d_total, d_stmp, s);
        fill_device_u32(d_deg, total_rows + 1, 0u, s);
        launch_fill(variant, grid_for(m), s, shape, ma, d_pts, sorted, d_flag, d_pos, m, rb, d_col, d_w, d_deg);
        TTG_LAUNCH_CHECK();
        exclusive_scan_u32_buf(d_deg, d_rp, total_rows + 1, nullptr, d_stmp, s);
    }
};

namespace {

void bknn_setup(BatchKnnState& b, u32 nclouds, const uint32_t* n, const uint32_t* dim, int kind, float p, u32 k) {
    check_metric(kind, p);
    if (nclouds == 0) fail(TT_EINVAL, "tt_batch_knn_create: no clouds");
    if (!n || !dim) fail(TT_EINVAL, "tt_batch_knn_create: null shapes");
    if (k == 0) fail(TT_EINVAL, "persistence: k must be >= 1");
    b.nclouds = nclouds;
    b.k = k;
    b.kind = kind;
    b.p = p;
    b.kmax = k <= 8 ? 8 : k <= 16 ? 16 : k <= 32 ? 32 : 64;
    b.variant = variant_of(kind, p);
    b.ma = metric_args(kind, p);
    b.n.assign(n, n + nclouds);
    b.dim.assign(dim, dim + nclouds);
    b.pt_off.assign(nclouds + 1, 0);
    b.row_base.assign(nclouds + 1, 0);
    b.list_off.assign(nclouds + 1, 0);
    u32 nmax = 1;
    for (u32 c = 0; c < nclouds; ++c) {
        if (b.n[c] && b.dim[c] == 0) fail(TT_EINVAL, "tt_batch_knn_create: dim must be >= 1");
        const u32 kc = b.n[c] >= 2 ? std::min(k, b.n[c] - 1) : 0u;
        b.pt_off[c + 1] = b.pt_off[c] + u64(b.n[c]) * b.dim[c];
        b.row_base[c + 1] = b.row_base[c] + b.n[c];
        b.list_off[c + 1] = b.list_off[c] + u64(b.n[c]) * kc;
        nmax = std::max(nmax, b.n[c]);
    }
    b.total_pts = b.pt_off[nclouds];
    b.total_rows = b.row_base[nclouds];
    b.total_lists = b.list_off[nclouds];
    b.m = 2 * b.total_lists;
    if (b.m >= (u64(1) << 32)) fail(TT_ELENGTH, "tt_batch_knn_create: more than 2^32 list entries in one batch");
    b.rb = std::max(1, bit_length(u64(nmax - 1)));
    b.cb = bit_length(u64(nclouds - 1));
    if (b.cb + 2 * b.rb > 64) fail(TT_ELENGTH, "tt_batch_knn_create: graph keys exceed 64 bits");

    // query blocks of the register-list scan, grouped by dimension
    // (cloud-major within a group); longer lists go to the selection path
    std::vector<u32> qc[kDimGroups], qf[kDimGroups];
    for (u32 c = 0; c < nclouds; ++c) {
        if (b.n[c] < 2) continue;
        const u32 kc = std::min(k, b.n[c] - 1);
        if (kc > 64) {
            if (u64(b.n[c]) * kc >= (u64(1) << 32)) fail(TT_ELENGTH, "knn: n * k must be below 2^32");
            b.large.push_back(c);
            continue;
        }
        const int g = dim_group(b.dim[c]);
        for (u32 q = 0; q < b.n[c]; q += QB) {
            qc[g].push_back(c);
            qf[g].push_back(q);
        }
    }
    std::vector<u32> qb_cloud, qb_first;
    b.gq.assign(kDimGroups + 1, 0);
    for (int g = 0; g < kDimGroups; ++g) {
        qb_cloud.insert(qb_cloud.end(), qc[g].begin(), qc[g].end());
        qb_first.insert(qb_first.end(), qf[g].begin(), qf[g].end());
        b.gq[g + 1] = qb_cloud.size();
    }
    std::vector<u32> row_cloud(std::max<u64>(b.total_rows, 1), 0);
    for (u32 c = 0; c < nclouds; ++c)
        std::fill(row_cloud.begin() + b.row_base[c], row_cloud.begin() + b.row_base[c + 1], c);

    Workspace& ws = b.ws;
    const u64 m1 = std::max<u64>(b.m, 1);
    b.d_pts = ws.get<float>(std::max<u64>(b.total_pts, 1));
    b.d_ids = ws.get<u32>(std::max<u64>(b.total_lists, 1));
    b.d_keys = ws.get<u64>(m1);
    b.d_alt = ws.get<u64>(m1);
    b.d_flag = ws.get<u32>(m1);
    b.d_pos = ws.get<u32>(m1);
    b.d_col = ws.get<u32>(m1);
    b.d_w = ws.get<float>(m1);
    b.d_deg = ws.get<u32>(b.total_rows + 1);
    b.d_rp = ws.get<u32>(b.total_rows + 1);
    b.d_total = ws.get<u32>(1);
    b.d_count = ws.get<u32>(1);
    b.d_sort = ws.get<u64>(radix_scratch_entries(m1));
    b.d_stmp = ws.get<u32>(std::max({scan_tmp_entries(m1),
                                     scan_tmp_entries(b.total_rows + 1), u64(1)}));
    b.d_n = ws.get<u32>(nclouds);
    b.d_dim = ws.get<u32>(nclouds);
    b.d_pt_off = ws.get<u64>(nclouds + 1);
    b.d_row_base = ws.get<u64>(nclouds + 1);
    b.d_list_off = ws.get<u64>(nclouds + 1);
    b.d_row_cloud = ws.get<u32>(row_cloud.size());
    b.d_qb_cloud = ws.get<u32>(std::max<size_t>(qb_cloud.size(), 1));
    b.d_qb_first = ws.get<u32>(std::max<size_t>(qb_first.size(), 1));
    {
        const u64 r1 = std::max<u64>(b.total_rows, 1);
        if (b.total_rows >= (u64(1) << 32)) fail(TT_ELENGTH, "tt_batch_knn_create: 2^32 points or more in one batch");
        b.codebits = u32(32 - b.cb);   // cloud bits and code bits share the top 32 of each key
        if (b.codebits > 23) b.codebits = 23;
        b.d_skeys = ws.get<u64>(r1);
        b.d_salt = ws.get<u64>(r1);
        b.d_ssort = ws.get<u64>(radix_scratch_entries(r1));
        b.d_xs = ws.get<float>(std::max<u64>(b.total_pts, 1));
        b.d_perm = ws.get<u32>(r1);
        b.d_ext = ws.get<u32>(u64(nclouds) * EXT_STRIDE);
        b.d_nrows = ws.get<u32>(1);
    }
    TTG_HIP(hipStreamCreateWithFlags(&b.stream, hipStreamNonBlocking));
    hipStream_t s = b.stream;
    const u32 count = u32(b.m);
    TTG_HIP(hipMemcpyAsync(b.d_count, &count, 4, hipMemcpyHostToDevice, s));
    const u32 nrows = u32(b.total_rows);
    TTG_HIP(hipMemcpyAsync(b.d_nrows, &nrows, 4, hipMemcpyHostToDevice, s));
    TTG_HIP(hipMemcpyAsync(b.d_n, b.n.data(), nclouds * 4, hipMemcpyHostToDevice, s));
    TTG_HIP(hipMemcpyAsync(b.d_dim, b.dim.data(), nclouds * 4, hipMemcpyHostToDevice, s));
    TTG_HIP(hipMemcpyAsync(b.d_pt_off, b.pt_off.data(), (nclouds + 1) * 8, hipMemcpyHostToDevice, s));
    TTG_HIP(hipMemcpyAsync(b.d_row_base, b.row_base.data(), (nclouds + 1) * 8, hipMemcpyHostToDevice, s));
    TTG_HIP(hipMemcpyAsync(b.d_list_off, b.list_off.data(), (nclouds + 1) * 8, hipMemcpyHostToDevice, s));
    TTG_HIP(hipMemcpyAsync(b.d_row_cloud, row_cloud.data(), row_cloud.size() * 4, hipMemcpyHostToDevice, s));
    if (!qb_cloud.empty()) {
        TTG_HIP(hipMemcpyAsync(b.d_qb_cloud, qb_cloud.data(), qb_cloud.size() * 4, hipMemcpyHostToDevice, s));
        TTG_HIP(hipMemcpyAsync(b.d_qb_first, qb_first.data(), qb_first.size() * 4, hipMemcpyHostToDevice, s));
    }
    TTG_HIP(hipStreamSynchronize(s));   // the host vectors above go out of scope
    b.shape = KnnBatchShape{b.d_n, b.d_dim, b.d_pt_off, b.d_row_base, b.d_list_off, b.d_row_cloud, k, b.total_rows};
    TTG_HIP(hipHostMalloc(reinterpret_cast<void**>(&b.h_pts), std::max<u64>(b.total_pts, 1) * 4,
                          hipHostMallocDefault));
    TTG_HIP(hipEventCreate(&b.t0));
    TTG_HIP(hipEventCreate(&b.t1));
}

void bknn_run(BatchKnnState& b, const float* pts, int use_graph) {
    hipStream_t s = b.stream;
    for (u64 i = 0; i < b.total_pts; ++i)
        if (!std::isfinite(pts[i])) fail(TT_EINVAL, "persistence: non-finite coordinate");
    if (b.total_pts) {
        std::memcpy(b.h_pts, pts, b.total_pts * 4);
        TTG_HIP(hipMemcpyAsync(b.d_pts, b.h_pts, b.total_pts * 4, hipMemcpyHostToDevice, s));
    }
    b.ws_large.release();   // the previous run has finished: its temporaries go
    TTG_HIP(hipEventRecord(b.t0, s));
    if (use_graph && b.large.empty()) {
        if (!b.exec) {
            TTG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            b.enqueue();
            TTG_HIP(hipStreamEndCapture(s, &b.graph));
            TTG_HIP(hipGraphInstantiate(&b.exec, b.graph, nullptr, nullptr, 0));
        }
        TTG_HIP(hipGraphLaunch(b.exec, s));
    } else {
        b.enqueue();
    }
    TTG_HIP(hipEventRecord(b.t1, s));
    b.h_rp.resize(b.total_rows + 1);
    TTG_HIP(hipMemcpyAsync(b.h_rp.data(), b.d_rp, (b.total_rows + 1) * 4, hipMemcpyDeviceToHost, s));
    TTG_HIP(hipStreamSynchronize(s));
    TTG_HIP(hipEventElapsedTime(&b.last_ms, b.t0, b.t1));
    const u64 nnz = b.h_rp[b.total_rows];
    b.h_col.resize(nnz);
    b.h_w.resize(nnz);
    if (nnz) {
        copy_to_host(b.h_col.data(), b.d_col, nnz * 4, s);
        copy_to_host(b.h_w.data(), b.d_w, nnz * 4, s);
    }
    b.ran = true;
}

}  // namespace

}  // namespace ttg

using namespace ttg;

struct tt_batch_knn {
    BatchKnnState st;
};

extern "C" {

tt_status tt_batch_knn_create(uint32_t nclouds, const uint32_t* n, const uint32_t* dim, int kind, float p,
                              uint32_t k, tt_batch_knn** out) {
    return guarded([&] {
        if (!out) fail(TT_EINVAL, "tt_batch_knn_create: null handle");
        *out = nullptr;
        auto* b = new tt_batch_knn;
        try {
            bknn_setup(b->st, nclouds, n, dim, kind, p, k);
        } catch (...) {
            delete b;
            throw;
        }
        *out = b;
    });
}

tt_status tt_batch_knn_run(tt_batch_knn* b, const float* pts, int use_graph) {
    return guarded([&] {
        if (!b) fail(TT_EINVAL, "tt_batch_knn_run: null handle");
        if (b->st.total_pts && !pts) fail(TT_EINVAL, "tt_batch_knn_run: null points");
        bknn_run(b->st, pts, use_graph);
    });
}

tt_status tt_batch_knn_cloud(tt_batch_knn* b, uint32_t cloud, const uint32_t** row_off, const uint32_t** col,
                             const float** w) {
    return guarded([&] {
        if (!b || cloud >= b->st.nclouds) fail(TT_EINVAL, "tt_batch_knn_cloud: bad cloud");
        BatchKnnState& s = b->st;
        if (!s.ran) fail(TT_EINVAL, "tt_batch_knn_cloud: run the batch first");
        if (row_off) *row_off = s.h_rp.data() + s.row_base[cloud];
        if (col) *col = s.h_col.data();
        if (w) *w = s.h_w.data();
    });
}

tt_status tt_batch_knn_lists(tt_batch_knn* b, uint32_t cloud, uint32_t* k, const uint32_t** ids) {
    return guarded([&] {
        if (!b || cloud >= b->st.nclouds) fail(TT_EINVAL, "tt_batch_knn_lists: bad cloud");
        BatchKnnState& s = b->st;
        if (!s.ran) fail(TT_EINVAL, "tt_batch_knn_lists: run the batch first");
        const u32 nc = s.n[cloud];
        const u32 kc = nc >= 2 ? std::min(s.k, nc - 1) : 0u;
        const u64 len = u64(nc) * kc;
        s.lists.resize(len);
        if (len) copy_to_host(s.lists.data(), s.d_ids + s.list_off[cloud], len * 4, s.stream);
        if (k) *k = kc;
        if (ids) *ids = s.lists.data();
    });
}

tt_status tt_batch_knn_info(tt_batch_knn* b, float* device_ms, uint64_t* nnz) {
    return guarded([&] {
        if (!b) fail(TT_EINVAL, "tt_batch_knn_info: null handle");
        if (device_ms) *device_ms = b->st.last_ms;
        if (nnz) *nnz = b->st.ran ? b->st.h_rp[b->st.total_rows] : 0;
    });
}

void tt_batch_knn_destroy(tt_batch_knn* b) { delete b; }

}  // extern "C"
