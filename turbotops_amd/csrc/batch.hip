// SPDX-License-Identifier: MIT
// batch.hip — the dense front end of many point clouds of uneven size and
// dimension in one launch sequence, recorded once into a HIP graph.
//
// Per cloud c it produces what persistence()'s dense branch needs
// (include/turbotops/turbotops.hpp:230-253, rips.hpp:391-416, :370-380,
// :685-849): the full distance matrix, the enclosing radius, and the kept
// edges of dimension 0 in filtration order that are NOT in an apparent pair.
// Apparent edges are never union-find tree edges (both other edges of their
// triangle precede them), so for H1 the host needs only the rest.
//
// All clouds share one set of buffers: points, matrices and rows are packed
// with per-cloud offsets. Kernels take fixed grids and read every
// data-dependent count from device memory, so the whole sequence can be
// replayed from a graph:
//   batch_dist     persistent, one launch per dimension group (2, 3, 8,
//                  other): blocks stride over the group's tile range; a
//                  per-tile cloud table names the cloud, and the tile runs the
//                  single-cloud tile body (tiles.hpp) specialised for the
//                  group's dimension
//   batch_radius   one workgroup per cloud: min of the row maxima
//   batch_count    one wave per global row: kept edges j < i
//   scan           global row offsets; cloud segments follow
//   batch_compact  keys (cloud, diameter bits, M_c - 1 - idx) in row-major
//                  order
//   batch_block_min  per row of every cloud, the minima of its 64-column blocks
//   batch_apparent 64 edges per wave (the two-phase apparent test of
//                  apparent.hpp), on the unsorted edges: the test does not
//                  depend on the order
//   batch_select   compaction of the non-apparent keys
//   counted radix sort of those keys over the device-side count: one sort
//                  orders every cloud by `precedes` and keeps clouds contiguous
//   batch_unpack   (idx, d) per non-apparent edge
// In the other edge order (tt_batch_set_order), for the boundary engine, the
// compaction keys hold the combinatorial index itself, the ready-witness test
// (batch_witness) replaces the apparent test and appends its flag below the
// key, and every kept edge is sorted and unpacked to its vertex pair.
// Host copies then run per cloud, each followed by an event, so the host
// reducer starts on a cloud as soon as its data has landed.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "apparent.hpp"
#include "frontend.hpp"
#include "tiles.hpp"

namespace ttg {

namespace {

struct BatchShape {
    u32 nclouds;
    const u32* n;
    const u32* dim;
    const u64* pt_off;
    const u64* mat_off;     // device matrices: rows padded to row_ld(n) floats
    const u64* row_base;    // nclouds + 1
    const u64* tile_base;   // nclouds + 1
    const u64* pair_base;   // nclouds + 1 (key capacity offsets)
    const u64* bmin_off;    // nclouds + 1 (row block minima offsets)
    const u32* tile_cloud;  // total_tiles: the cloud of each tile
    const u32* row_cloud;   // total_rows: the cloud of each row
    u64 total_tiles, total_rows;
};

// Row stride of a device matrix: n rounded up to 32 floats, so that every
// row starts on a 128-byte line and the tiles' 256-byte row segments are
// whole lines (a segment that straddles lines halves the write rate). The
// host copy is unpadded.
__host__ __device__ __forceinline__ u32 row_ld(u32 n) { return (n + 31u) & ~31u; }

// Dimension groups of the tile list: tiles are numbered group by group
// (clouds of dimension 2, then 3, then 8, then every other dimension), cloud
// by cloud within a group, so each group is one contiguous tile range and one
// launch of a kernel specialised for its dimension (0: runtime dimension).
// One kernel over every dimension holds the registers of its heaviest branch
// (the d = 8 rows) for all of them and spills SGPRs.
constexpr int kDimGroups = 4;
constexpr u32 kGroupDim[kDimGroups] = {2, 3, 8, 0};
inline int dim_group(u32 d) { return d == 2 ? 0 : d == 3 ? 1 : d == 8 ? 2 : 3; }

template <int V, int RPW, int DIM>
__global__ __launch_bounds__(64 * (64 / RPW)) void batch_dist(BatchShape b, MetricArgs ma,
                                                              const float* __restrict__ pts,
                                                              float* __restrict__ mats,
                                                              u32* __restrict__ rowmax, u64 t_begin,
                                                              u64 t_end) {
    __shared__ TileSmem<RPW> sm;
    // Tiles in order: cloud by cloud, row-major lower triangle within one.
    // (A scattered order, t * A mod T, meant to spread the row-maximum
    // atomics of small clouds over more cache lines, measured 6% slower on
    // the 512-cloud set: 3.01 against 2.83 ms.)
    for (u64 t = t_begin + blockIdx.x; t < t_end; t += gridDim.x) {
        const u32 c = __builtin_amdgcn_readfirstlane(b.tile_cloud[t]);
        const u32 n = b.n[c];
        const float* x = pts + b.pt_off[c];
        float* mat = mats + b.mat_off[c];
        u32* rm = rowmax + b.row_base[c];
        u32 bi, bj;
        lower_tile_of(t - b.tile_base[c], bi, bj);
        dist_tile<V, DIM, TT_LAYOUT_FULL, RPW>(x, n, DIM ? u32(DIM) : b.dim[c], ma, bi, bj, mat, rm, sm,
                                               row_ld(n));
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void batch_radius(BatchShape b, const u32* __restrict__ rowmax,
                                                    float user_thr, float* __restrict__ thr) {
    __shared__ u32 red[4];
    const u32 c = blockIdx.x;
    const u32 n = b.n[c];
    const u32* rm = rowmax + b.row_base[c];
    u32 m = 0x7f800000u;
    for (u32 t = threadIdx.x; t < n; t += 256) m = min(m, rm[t]);
    for (int off = 32; off > 0; off >>= 1) m = min(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        const u32 r = min(min(red[0], red[1]), min(red[2], red[3]));
        thr[c] = isinf(user_thr) ? __uint_as_float(r) : user_thr;
    }
}

constexpr u32 BCH = 8;   // row chunks of 64 columns in flight per wave (batch_count, batch_compact)

__global__ __launch_bounds__(256) void batch_count(BatchShape b, const float* __restrict__ mats,
                                                   const float* __restrict__ thr, u32* __restrict__ cnt) {
    const u32 lane = threadIdx.x & 63;
    const u64 gw = (blockIdx.x * 256ull + threadIdx.x) >> 6;
    const u64 nw = (u64(gridDim.x) * 256) >> 6;
    for (u64 r = gw; r < b.total_rows; r += nw) {
        const u32 c = b.row_cloud[r];
        const u32 n = b.n[c];
        const u32 i = u32(r - b.row_base[c]);
        const float* row = mats + b.mat_off[c] + u64(i) * row_ld(n);
        const float th = thr[c];
        u32 k = 0;
        // BCH loads of 64 columns issued before any is used (one per step
        // left the stage latency-bound)
        for (u32 j0 = 0; j0 < i; j0 += 64 * BCH) {
            float d[BCH];
#pragma unroll
            for (u32 u = 0; u < BCH; ++u) {
                const u32 j = j0 + 64 * u + lane;
                d[u] = j < i ? row[j] : 0.0f;
            }
#pragma unroll
            for (u32 u = 0; u < BCH; ++u) k += (j0 + 64 * u + lane < i && d[u] <= th) ? 1u : 0u;
        }
#pragma unroll
        for (u32 off = 32; off > 0; off >>= 1) k += __shfl_xor(k, off);
        if (lane == 0) cnt[r] = k;
    }
}

__global__ void batch_segments(BatchShape b, const u32* __restrict__ off, u32* __restrict__ seg) {
    for (u32 c = threadIdx.x; c <= b.nclouds; c += blockDim.x) seg[c] = off[b.row_base[c]];
}

__global__ __launch_bounds__(256) void batch_compact(BatchShape b, const float* __restrict__ mats,
                                                     const float* __restrict__ thr,
                                                     const u32* __restrict__ off, int ib, bool comb,
                                                     u64* __restrict__ keys) {
    const u32 lane = threadIdx.x & 63;
    const u64 gw = (blockIdx.x * 256ull + threadIdx.x) >> 6;
    const u64 nw = (u64(gridDim.x) * 256) >> 6;
    for (u64 r = gw; r < b.total_rows; r += nw) {
        const u32 c = b.row_cloud[r];
        const u32 n = b.n[c];
        const u32 i = u32(r - b.row_base[c]);
        const float* row = mats + b.mat_off[c] + u64(i) * row_ld(n);
        const float th = thr[c];
        const u64 M = u64(n) * (n - 1) / 2;
        const u64 cloud_bits = u64(c) << (31 + ib);
        u64 o = off[r];
        // comb: the cloud's keys in row-major order, which is ascending
        // index; otherwise mirrored within the cloud's segment, descending
        // index. Either way the order the sort must keep among equal
        // diameters is already there, so the sort runs over the (cloud,
        // diameter) bits alone and keeps it by stability.
        const u64 mirror = u64(off[b.row_base[c]]) + off[b.row_base[c + 1]] - 1;
        const u64 rowbase = u64(i) * (i - 1) / 2;
        for (u32 j0 = 0; j0 < i; j0 += 64 * BCH) {
            float d[BCH];
#pragma unroll
            for (u32 u = 0; u < BCH; ++u) {
                const u32 j = j0 + 64 * u + lane;
                d[u] = j < i ? row[j] : 0.0f;
            }
#pragma unroll
            for (u32 u = 0; u < BCH; ++u) {   // chunks in column order
                const u32 j = j0 + 64 * u + lane;
                const bool keep = j < i && d[u] <= th;
                const u64 bal = __ballot(keep);
                if (keep) {
                    const u32 rk = __builtin_amdgcn_mbcnt_hi(u32(bal >> 32), __builtin_amdgcn_mbcnt_lo(u32(bal), 0u));
                    const u64 idx = rowbase + j;
                    const u64 at = comb ? o + rk : mirror - (o + rk);
                    keys[at] = cloud_bits | (u64(__float_as_uint(d[u])) << ib) | (comb ? idx : M - 1 - idx);
                }
                o += __popcll(bal);
            }
        }
    }
}

// Candidates the lanes walk together before the grouped search, in the batch's
// apparent and witness scans. The C4 clouds are unions of circles stored by
// component, so a quarter of their edges find no cofacet among the top
// candidates and go to the grouped search anyway: walking 16 instead of 64
// measured 137 against 149 ms (witness order) and 101 against 110 ms
// (cohomology order) for the 512 clouds; 8 and 32 were slower. The
// single-cloud scans keep TOPK (64), best on uniform and Gaussian clouds.
constexpr u32 kBatchTopk = 16;

__device__ __forceinline__ void decode_edge_b(u64 idx, u32& i, u32& j) {
    u64 v = u64((1.0 + __builtin_sqrt(1.0 + 8.0 * double(idx))) * 0.5);
    while (v * (v - 1) / 2 > idx) --v;
    while ((v + 1) * v / 2 <= idx) ++v;
    i = u32(v);
    j = u32(idx - v * (v - 1) / 2);
}

// Apparent test per sorted edge (one wave64 each) and the keep flag of the
// non-apparent ones.
// Block minima of every row of every cloud's matrix, one wave per row.
__global__ __launch_bounds__(256) void batch_block_min(BatchShape b, const float* __restrict__ mats,
                                                       float* __restrict__ bmins) {
    const u32 lane = threadIdx.x & 63;
    const u64 gw = (blockIdx.x * 256ull + threadIdx.x) >> 6;
    const u64 nw = (u64(gridDim.x) * 256) >> 6;
    for (u64 r = gw; r < b.total_rows; r += nw) {
        const u32 c = b.row_cloud[r];
        const u32 n = b.n[c];
        const u32 i = u32(r - b.row_base[c]);
        const u32 nb = (n + BLK - 1) / BLK;
        block_min_row(mats + b.mat_off[c] + u64(i) * row_ld(n), n, nb, bmins + b.bmin_off[c] + u64(i) * nb,
                      lane);
    }
}

__global__ __launch_bounds__(256) void batch_apparent(BatchShape b, const float* __restrict__ mats,
                                                      const float* __restrict__ bmins,
                                                      const u64* __restrict__ keys,
                                                      const u32* __restrict__ total, int ib,
                                                      u32* __restrict__ keep) {
    const u32 lane = threadIdx.x & 63;
    const u64 gw = (blockIdx.x * 256ull + threadIdx.x) >> 6;
    const u64 nw = (u64(gridDim.x) * 256) >> 6;
    const u64 m = *total;
    const u64 imask = (u64(1) << ib) - 1;
    for (u64 e0 = gw * 64; e0 < m; e0 += nw * 64) {
        const u64 e = e0 + lane;
        const bool valid = e < m;
        u32 i = 1, j = 0, n = 2, nb = 1;
        float d = 0.0f;
        const float* mat = mats;
        const float* bm = bmins;
        if (valid) {
            const u64 key = keys[e];
            const u32 c = u32(key >> (31 + ib));
            d = __uint_as_float(u32(key >> ib) & 0x7fffffffu);
            n = b.n[c];
            nb = (n + BLK - 1) / BLK;
            const u64 M = u64(n) * (n - 1) / 2;
            decode_edge_b(M - 1 - (key & imask), i, j);
            mat = mats + b.mat_off[c];
            bm = bmins + b.bmin_off[c];
        }
        const bool ap = edges_apparent_lanes<kBatchTopk>(mat, bm, n, row_ld(n), nb, valid, i, j, d, lane);
        if (valid) keep[e] = ap ? 0u : 1u;
    }
}

// The keys of the non-apparent edges, in compaction order.
__global__ void batch_select(const u64* __restrict__ keys, const u32* __restrict__ keep,
                             const u32* __restrict__ pos, const u32* __restrict__ total,
                             u64* __restrict__ out) {
    const u64 m = *total;
    for (u64 e = blockIdx.x * 256ull + threadIdx.x; e < m; e += u64(gridDim.x) * 256)
        if (keep[e]) out[pos[e]] = keys[e];
}

// Sorted keys of the non-apparent edges -> (idx, d).
__global__ void batch_unpack(const u64* __restrict__ keys, const u32* __restrict__ kept, int ib,
                             const u32* __restrict__ n_of, u64* __restrict__ out_idx,
                             float* __restrict__ out_d) {
    const u64 m = *kept;
    const u64 imask = (u64(1) << ib) - 1;
    for (u64 e = blockIdx.x * 256ull + threadIdx.x; e < m; e += u64(gridDim.x) * 256) {
        const u64 key = keys[e];
        const u32 c = u32(key >> (31 + ib));
        const u64 n = n_of[c];
        const u64 M = n * (n - 1) / 2;
        out_idx[e] = M - 1 - (key & imask);
        out_d[e] = __uint_as_float(u32(key >> ib) & 0x7fffffffu);
    }
}

// The apparent flag of every compacted coboundary-order key (cloud,
// diameter bits, M - 1 - idx) as its new lowest bit: every kept edge goes on
// in the cohomology engine's order with its flag (maxdim 0 or >= 2, where
// the next dimension is assembled from all edges).
__global__ __launch_bounds__(256) void batch_apparent_tag(BatchShape b, const float* __restrict__ mats,
                                                          const float* __restrict__ bmins, u64* __restrict__ keys,
                                                          const u32* __restrict__ total, int ib) {
    const u32 lane = threadIdx.x & 63;
    const u64 gw = (blockIdx.x * 256ull + threadIdx.x) >> 6;
    const u64 nw = (u64(gridDim.x) * 256) >> 6;
    const u64 m = *total;
    const u64 imask = (u64(1) << ib) - 1;
    for (u64 e0 = gw * 64; e0 < m; e0 += nw * 64) {
        const u64 e = e0 + lane;
        const bool valid = e < m;
        u32 i = 1, j = 0, n = 2, nb = 1;
        float d = 0.0f;
        u64 key = 0;
        const float* mat = mats;
        const float* bm = bmins;
        if (valid) {
            key = keys[e];
            const u32 c = u32(key >> (31 + ib));
            d = __uint_as_float(u32(key >> ib) & 0x7fffffffu);
            n = b.n[c];
            nb = (n + BLK - 1) / BLK;
            const u64 M = u64(n) * (n - 1) / 2;
            decode_edge_b(M - 1 - (key & imask), i, j);
            mat = mats + b.mat_off[c];
            bm = bmins + b.bmin_off[c];
        }
        const bool ap = edges_apparent_lanes<kBatchTopk>(mat, bm, n, row_ld(n), nb, valid, i, j, d, lane);
        if (valid) keys[e] = (key << 1) | (ap ? 1u : 0u);
    }
}

// Sorted tagged keys -> (idx, d, apparent flag) per kept edge.
__global__ void batch_unpack_tagged(const u64* __restrict__ keys, const u32* __restrict__ total, int ib,
                                    const u32* __restrict__ n_of, u64* __restrict__ out_idx,
                                    float* __restrict__ out_d, u8* __restrict__ out_ap) {
    const u64 m = *total;
    const u64 imask = (u64(1) << ib) - 1;
    for (u64 e = blockIdx.x * 256ull + threadIdx.x; e < m; e += u64(gridDim.x) * 256) {
        const u64 raw = keys[e];
        const u64 key = raw >> 1;
        const u32 c = u32(key >> (31 + ib));
        const u64 n = n_of[c];
        const u64 M = n * (n - 1) / 2;
        out_idx[e] = M - 1 - (key & imask);
        out_d[e] = __uint_as_float(u32(key >> ib) & 0x7fffffffu);
        out_ap[e] = u8(raw & 1);
    }
}

// The boundary engine's ready-witness flag of every compacted comb-order key
// (cloud, diameter bits, comb index), 64 edges per wave (apparent.hpp); the
// flag becomes the key's lowest bit.
__global__ __launch_bounds__(256) void batch_witness(BatchShape b, const float* __restrict__ mats,
                                                     const float* __restrict__ bmins, u64* __restrict__ keys,
                                                     const u32* __restrict__ total, int ib) {
    const u32 lane = threadIdx.x & 63;
    const u64 gw = (blockIdx.x * 256ull + threadIdx.x) >> 6;
    const u64 nw = (u64(gridDim.x) * 256) >> 6;
    const u64 m = *total;
    const u64 imask = (u64(1) << ib) - 1;
    for (u64 e0 = gw * 64; e0 < m; e0 += nw * 64) {
        const u64 e = e0 + lane;
        const bool valid = e < m;
        u32 i = 1, j = 0, n = 2, nb = 1;
        float d = 0.0f;
        u64 key = 0;
        const float* mat = mats;
        const float* bm = bmins;
        if (valid) {
            key = keys[e];
            const u32 c = u32(key >> (31 + ib));
            d = __uint_as_float(u32(key >> ib) & 0x7fffffffu);
            n = b.n[c];
            nb = (n + BLK - 1) / BLK;
            decode_edge_b(key & imask, i, j);
            mat = mats + b.mat_off[c];
            bm = bmins + b.bmin_off[c];
        }
        const bool wf = edges_apparent_lanes<kBatchTopk, true>(mat, bm, n, row_ld(n), nb, valid, i, j, d, lane);
        if (valid) keys[e] = (key << 1) | (wf ? 1u : 0u);
    }
}

// Sorted keys of every kept edge in (diameter, combinatorial index) order ->
// (lo | hi << 32, d): the vertex pair reads as two u32 (lo, hi) in memory.
// fb = 1: the key carries the witness flag in bit 0, which goes to bit 31 of hi.
__global__ void batch_unpack_comb(const u64* __restrict__ keys, const u32* __restrict__ total, int ib, int fb,
                                  u64* __restrict__ out_pair, float* __restrict__ out_d) {
    const u64 m = *total;
    const u64 imask = (u64(1) << ib) - 1;
    for (u64 e = blockIdx.x * 256ull + threadIdx.x; e < m; e += u64(gridDim.x) * 256) {
        const u64 raw = keys[e];
        const u64 key = raw >> fb;
        const u32 flag = fb ? u32(raw & 1) : 0u;
        u32 hi, lo;
        decode_edge_b(key & imask, hi, lo);
        out_pair[e] = u64(lo) | (u64(hi | (flag << 31)) << 32);
        out_d[e] = __uint_as_float(u32(key >> ib) & 0x7fffffffu);
    }
}

__global__ void copy_u32(const u32* __restrict__ src, u32 count, u32* __restrict__ dst) {
    for (u32 t = threadIdx.x; t < count; t += blockDim.x) dst[t] = src[t];
}

// per cloud: start of its non-apparent edges in the sorted list = number of
// non-apparent edges before its first compacted edge (compaction and sort are
// both cloud-major)
__global__ void batch_kept_segments(const u32* __restrict__ seg, const u32* __restrict__ pos,
                                    const u32* __restrict__ keep, const u32* __restrict__ total,
                                    u32 nclouds, u32* __restrict__ kseg) {
    for (u32 c = threadIdx.x; c <= nclouds; c += blockDim.x) {
        const u32 s = seg[c];
        const u32 m = *total;
        kseg[c] = s < m ? pos[s] : (m ? pos[m - 1] + keep[m - 1] : 0u);
    }
}

template <int V, int RPW>
void launch_batch_dist_group(int g, const u64* gt, hipStream_t s, const BatchShape& b, MetricArgs ma,
                             const float* pts, float* mats, u32* rowmax) {
    constexpr u32 threads = 64 * (64 / RPW);
    const u64 t0 = gt[g], t1 = gt[g + 1];
    if (t1 <= t0) return;
    const u32 grid = u32(t1 - t0);   // one tile per workgroup
    switch (kGroupDim[g]) {
        case 2: batch_dist<V, RPW, 2><<<grid, threads, 0, s>>>(b, ma, pts, mats, rowmax, t0, t1); break;
        case 3: batch_dist<V, RPW, 3><<<grid, threads, 0, s>>>(b, ma, pts, mats, rowmax, t0, t1); break;
        case 8: batch_dist<V, RPW, 8><<<grid, threads, 0, s>>>(b, ma, pts, mats, rowmax, t0, t1); break;
        default: batch_dist<V, RPW, 0><<<grid, threads, 0, s>>>(b, ma, pts, mats, rowmax, t0, t1); break;
    }
}

template <int V>
void launch_batch_dist_rpw(int g, int rpw, const u64* gt, hipStream_t s, const BatchShape& b, MetricArgs ma,
                           const float* pts, float* mats, u32* rowmax) {
    switch (rpw) {
        case 8: launch_batch_dist_group<V, 8>(g, gt, s, b, ma, pts, mats, rowmax); break;
        case 32: launch_batch_dist_group<V, 32>(g, gt, s, b, ma, pts, mats, rowmax); break;
        default: launch_batch_dist_group<V, 16>(g, gt, s, b, ma, pts, mats, rowmax); break;
    }
}

// One dimension group g (its tile range gt[g], gt[g + 1]) with `rpw` rows per
// wave; g < 0 launches every group, group h with rpw_g[h].
void launch_batch_dist_v(MetricVariant v, int g, int rpw, const int* rpw_g, const u64* gt, hipStream_t s,
                         const BatchShape& b, MetricArgs ma, const float* pts, float* mats, u32* rowmax) {
    for (int h = (g < 0 ? 0 : g); h < (g < 0 ? kDimGroups : g + 1); ++h) {
        const int r = g < 0 ? rpw_g[h] : rpw;
        switch (v) {
            case MV_L1: launch_batch_dist_rpw<MV_L1>(h, r, gt, s, b, ma, pts, mats, rowmax); break;
            case MV_L2: launch_batch_dist_rpw<MV_L2>(h, r, gt, s, b, ma, pts, mats, rowmax); break;
            case MV_LINF: launch_batch_dist_rpw<MV_LINF>(h, r, gt, s, b, ma, pts, mats, rowmax); break;
            case MV_LP_INT: launch_batch_dist_rpw<MV_LP_INT>(h, r, gt, s, b, ma, pts, mats, rowmax); break;
            default: launch_batch_dist_rpw<MV_LP_POW>(h, r, gt, s, b, ma, pts, mats, rowmax); break;
        }
    }
}

// Autotuned tile shape per (metric variant, dimension set) and dimension
// group: the fastest rows-per-wave of {8, 16, 32} for each group's launch,
// measured on the batch itself.
struct TileShapes {
    int rpw[kDimGroups];
};
std::mutex g_tune_mu;
std::map<std::pair<int, u64>, TileShapes> g_tune;

}  // namespace

struct BatchState {
    u32 nclouds = 0;
    int kind = 1;
    float p = 2.0f, threshold = INFINITY;
    MetricVariant variant = MV_L2;
    MetricArgs ma{};
    int rpw_g[kDimGroups] = {16, 16, 16, 16};   // rows per wave of each dimension group's tiles
    int ib = 1, cb = 0;
    bool tuned = false;
    bool comb = false;   // hand over every kept edge in (diameter, comb index) order
    bool witness = false;   // ... with the ready-witness flags (asked for, and the keys have a bit to spare)
    bool witness_fits = false;
    bool tagged = false;   // every kept edge in the coboundary order with its apparent flag
    std::vector<u32> n, dim;
    std::vector<u64> pt_off, mat_off, dmat_off, row_base, tile_base, pair_base, bmin_off;
    u64 total_pts = 0, total_mat = 0, total_rows = 0, total_tiles = 0, capacity = 0, h_cap = 0;
    u64 group_tiles[kDimGroups + 1] = {};   // tile range of each dimension group
    // device
    Workspace ws;
    float *d_pts = nullptr, *d_mats = nullptr, *d_thr = nullptr, *d_out_d = nullptr, *d_bmin = nullptr;
    u32 *d_rowmax = nullptr, *d_cnt = nullptr, *d_total = nullptr, *d_seg = nullptr, *d_keep = nullptr,
        *d_pos = nullptr, *d_kept = nullptr, *d_kseg = nullptr, *d_stmp = nullptr,
        *d_n = nullptr, *d_dim = nullptr;
    u8* d_out_ap = nullptr;
    u64 *d_keys = nullptr, *d_alt = nullptr, *d_out_idx = nullptr, *d_pt_off = nullptr, *d_mat_off = nullptr,
        *d_row_base = nullptr, *d_tile_base = nullptr, *d_pair_base = nullptr, *d_bmin_off = nullptr;
    u64* d_sort = nullptr;   // radix sort scratch (radix_scratch_entries)
    u32 *d_tile_cloud = nullptr, *d_row_cloud = nullptr;
    BatchShape shape{};
    // host (pinned)
    float* h_pts = nullptr;
    float* h_mats = nullptr;
    u64* h_idx = nullptr;
    float* h_d = nullptr;
    u8* h_ap = nullptr;
    std::vector<u32> seg, kseg;
    std::vector<float> thr;
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<hipEvent_t> ev;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    float last_ms = 0.0f;
    ~BatchState() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (auto e : ev) (void)hipEventDestroy(e);
        if (t0) (void)hipEventDestroy(t0);
        if (t1) (void)hipEventDestroy(t1);
        if (stream) (void)hipStreamDestroy(stream);
        if (h_pts) (void)hipHostFree(h_pts);
        if (h_mats) (void)hipHostFree(h_mats);
        if (h_idx) (void)hipHostFree(h_idx);
        if (h_d) (void)hipHostFree(h_d);
        if (h_ap) (void)hipHostFree(h_ap);
    }

    // the device part of one batch (recorded into the graph)
    // rows per wave of the group with the most tiles (reported)
    int rpw_main() const {
        int g0 = 0;
        for (int g = 1; g < kDimGroups; ++g)
            if (group_tiles[g + 1] - group_tiles[g] > group_tiles[g0 + 1] - group_tiles[g0]) g0 = g;
        return rpw_g[g0];
    }

    void enqueue() {
        hipStream_t s = stream;
        fill_device_u32(d_rowmax, std::max<u64>(total_rows, 1), 0u, s);
        if (total_tiles)
            launch_batch_dist_v(variant, -1, 0, rpw_g, group_tiles, s, shape, ma, d_pts, d_mats, d_rowmax);
        TTG_LAUNCH_CHECK();
        batch_radius<<<nclouds, 256, 0, s>>>(shape, d_rowmax, threshold, d_thr);
        fill_device_u32(d_cnt, total_rows + 1, 0u, s);
        const u32 gw = u32(std::min<u64>(std::max<u64>((total_rows + 3) / 4, 1), 16384));
        batch_count<<<gw, 256, 0, s>>>(shape, d_mats, d_thr, d_cnt);
        exclusive_scan_u32_buf(d_cnt, d_cnt, total_rows + 1, d_total, d_stmp, s);
        batch_segments<<<1, 256, 0, s>>>(shape, d_cnt, d_seg);
        batch_compact<<<gw, 256, 0, s>>>(shape, d_mats, d_thr, d_cnt, ib, comb, d_keys);
        TTG_LAUNCH_CHECK();
        const u32 ge = u32(std::min<u64>(std::max<u64>(capacity / 4, 1), 65536));
        if (comb) {
            // every kept edge, sorted by (cloud, diameter, comb index); the
            // cloud segments are those of the compaction
            const int fb = witness ? 1 : 0;
            if (witness) {
                batch_block_min<<<gw, 256, 0, s>>>(shape, d_mats, d_bmin);
                batch_witness<<<ge, 256, 0, s>>>(shape, d_mats, d_bmin, d_keys, d_total, ib);
            }
            const u64* sorted =
                radix_sort_u64_counted(d_keys, capacity, d_total, ib + fb, cb + 31 + ib + fb, d_alt, d_sort, s);
            batch_unpack_comb<<<ge, 256, 0, s>>>(sorted, d_total, ib, fb, d_out_idx, d_out_d);
            copy_u32<<<1, 256, 0, s>>>(d_seg, nclouds + 1, d_kseg);
            TTG_LAUNCH_CHECK();
            return;
        }
        if (tagged) {
            batch_block_min<<<gw, 256, 0, s>>>(shape, d_mats, d_bmin);
            batch_apparent_tag<<<ge, 256, 0, s>>>(shape, d_mats, d_bmin, d_keys, d_total, ib);
            const u64* sorted =
                radix_sort_u64_counted(d_keys, capacity, d_total, ib + 1, cb + 31 + ib + 1, d_alt, d_sort, s);
            batch_unpack_tagged<<<ge, 256, 0, s>>>(sorted, d_total, ib, d_n, d_out_idx, d_out_d, d_out_ap);
            copy_u32<<<1, 256, 0, s>>>(d_seg, nclouds + 1, d_kseg);
            TTG_LAUNCH_CHECK();
            return;
        }
        // the apparent test does not depend on the order, so it runs on the
        // compacted (row-major) edges and only the rest are sorted
        fill_device_u32(d_keep, capacity, 0u, s);   // slots past the count stay 0
        batch_block_min<<<gw, 256, 0, s>>>(shape, d_mats, d_bmin);
        batch_apparent<<<ge, 256, 0, s>>>(shape, d_mats, d_bmin, d_keys, d_total, ib, d_keep);
        exclusive_scan_u32_buf(d_keep, d_pos, capacity, d_kept, d_stmp, s);
        batch_select<<<ge, 256, 0, s>>>(d_keys, d_keep, d_pos, d_total, d_alt);
        batch_kept_segments<<<1, 256, 0, s>>>(d_seg, d_pos, d_keep, d_total, nclouds, d_kseg);
        TTG_LAUNCH_CHECK();
        const u64* sorted = radix_sort_u64_counted(d_alt, capacity, d_kept, ib, cb + 31 + ib, d_keys, d_sort, s);
        batch_unpack<<<ge, 256, 0, s>>>(sorted, d_kept, ib, d_n, d_out_idx, d_out_d);
        TTG_LAUNCH_CHECK();
    }
};

}  // namespace ttg

using namespace ttg;

struct tt_batch {
    BatchState st;
};

namespace {

void batch_setup(BatchState& b, u32 nclouds, const uint32_t* n, const uint32_t* dim, int kind, float p,
                 float threshold) {
    check_metric(kind, p);
    if (nclouds == 0) fail(TT_EINVAL, "tt_batch_create: no clouds");
    if (!n || !dim) fail(TT_EINVAL, "tt_batch_create: null shapes");
    if (std::isnan(threshold) || threshold < 0.0f)
        fail(TT_EINVAL, "persistence: threshold must be non-negative or +inf");
    b.nclouds = nclouds;
    b.kind = kind;
    b.p = p;
    b.threshold = threshold;
    b.variant = variant_of(kind, p);
    b.ma = metric_args(kind, p);
    b.n.assign(n, n + nclouds);
    b.dim.assign(dim, dim + nclouds);
    b.pt_off.assign(nclouds + 1, 0);
    b.mat_off.assign(nclouds + 1, 0);
    b.dmat_off.assign(nclouds + 1, 0);
    b.row_base.assign(nclouds + 1, 0);
    b.tile_base.assign(nclouds + 1, 0);
    b.pair_base.assign(nclouds + 1, 0);
    b.bmin_off.assign(nclouds + 1, 0);
    u64 maxM = 1;
    std::vector<u64> ntiles(nclouds);
    for (u32 c = 0; c < nclouds; ++c) {
        if (b.n[c] && b.dim[c] == 0) fail(TT_EINVAL, "tt_batch_create: dim must be >= 1");
        if (b.n[c] > 92682) fail(TT_ELENGTH, "tt_batch_create: a cloud exceeds 92682 points");
        const u64 nb = (u64(b.n[c]) + DTILE - 1) / DTILE;
        const u64 M = u64(b.n[c]) * (b.n[c] ? b.n[c] - 1 : 0) / 2;
        maxM = std::max<u64>(maxM, M);
        b.pt_off[c + 1] = b.pt_off[c] + u64(b.n[c]) * b.dim[c];
        b.mat_off[c + 1] = b.mat_off[c] + u64(b.n[c]) * b.n[c];
        b.dmat_off[c + 1] = b.dmat_off[c] + u64(b.n[c]) * row_ld(b.n[c]);
        b.row_base[c + 1] = b.row_base[c] + b.n[c];
        ntiles[c] = nb * (nb + 1) / 2;
        b.pair_base[c + 1] = b.pair_base[c] + M;
        b.bmin_off[c + 1] = b.bmin_off[c] + u64(b.n[c]) * ((u64(b.n[c]) + BLK - 1) / BLK);
    }
    b.total_pts = b.pt_off[nclouds];
    b.total_mat = b.mat_off[nclouds];
    b.total_rows = b.row_base[nclouds];
    // tile_base[c]: the first tile of cloud c, numbered group by group
    {
        u64 gcount[kDimGroups] = {};
        for (u32 c = 0; c < nclouds; ++c) gcount[dim_group(b.dim[c])] += ntiles[c];
        for (int g = 0; g < kDimGroups; ++g) b.group_tiles[g + 1] = b.group_tiles[g] + gcount[g];
        u64 next[kDimGroups];
        for (int g = 0; g < kDimGroups; ++g) next[g] = b.group_tiles[g];
        for (u32 c = 0; c < nclouds; ++c) {
            u64& at = next[dim_group(b.dim[c])];
            b.tile_base[c] = at;
            at += ntiles[c];
        }
        b.tile_base[nclouds] = b.group_tiles[kDimGroups];
    }
    b.total_tiles = b.group_tiles[kDimGroups];
    b.capacity = std::max<u64>(b.pair_base[nclouds], 1);
    if (b.capacity >= (u64(1) << 32)) fail(TT_ELENGTH, "tt_batch_create: more than 2^32 vertex pairs in one batch");
    b.ib = std::max(1, bit_length(maxM - 1));
    b.cb = bit_length(u64(nclouds - 1));
    if (b.cb + 31 + b.ib > 64) fail(TT_ELENGTH, "tt_batch_create: edge keys exceed 64 bits");
    b.witness_fits = b.cb + 31 + b.ib + 1 <= 64;

    Workspace& ws = b.ws;
    b.d_pts = ws.get<float>(std::max<u64>(b.total_pts, 1));
    b.d_mats = ws.get<float>(std::max<u64>(b.dmat_off[nclouds], 1));
    b.d_rowmax = ws.get<u32>(std::max<u64>(b.total_rows, 1));
    b.d_thr = ws.get<float>(nclouds);
    b.d_cnt = ws.get<u32>(b.total_rows + 1);
    b.d_total = ws.get<u32>(1);
    b.d_seg = ws.get<u32>(nclouds + 1);
    b.d_kseg = ws.get<u32>(nclouds + 1);
    b.d_keys = ws.get<u64>(b.capacity);
    b.d_alt = ws.get<u64>(b.capacity);
    b.d_keep = ws.get<u32>(b.capacity);
    b.d_pos = ws.get<u32>(b.capacity);
    b.d_kept = ws.get<u32>(1);
    b.d_out_idx = ws.get<u64>(b.capacity);
    b.d_out_d = ws.get<float>(b.capacity);
    b.d_sort = ws.get<u64>(radix_scratch_entries(b.capacity));
    const u64 stmp = std::max({scan_tmp_entries(b.capacity),
                               scan_tmp_entries(b.total_rows + 1), u64(1)});
    b.d_stmp = ws.get<u32>(stmp);
    b.d_n = ws.get<u32>(nclouds);
    b.d_dim = ws.get<u32>(nclouds);
    b.d_pt_off = ws.get<u64>(nclouds + 1);
    b.d_mat_off = ws.get<u64>(nclouds + 1);
    b.d_row_base = ws.get<u64>(nclouds + 1);
    b.d_tile_base = ws.get<u64>(nclouds + 1);
    b.d_pair_base = ws.get<u64>(nclouds + 1);
    b.d_bmin_off = ws.get<u64>(nclouds + 1);
    b.d_tile_cloud = ws.get<u32>(std::max<u64>(b.total_tiles, 1));
    b.d_row_cloud = ws.get<u32>(std::max<u64>(b.total_rows, 1));
    b.d_bmin = ws.get<float>(std::max<u64>(b.bmin_off[nclouds], 1));
    TTG_HIP(hipStreamCreateWithFlags(&b.stream, hipStreamNonBlocking));
    hipStream_t s = b.stream;
    TTG_HIP(hipMemcpyAsync(b.d_n, b.n.data(), nclouds * 4, hipMemcpyHostToDevice, s));
    TTG_HIP(hipMemcpyAsync(b.d_dim, b.dim.data(), nclouds * 4, hipMemcpyHostToDevice, s));
    TTG_HIP(hipMemcpyAsync(b.d_pt_off, b.pt_off.data(), (nclouds + 1) * 8, hipMemcpyHostToDevice, s));
    TTG_HIP(hipMemcpyAsync(b.d_mat_off, b.dmat_off.data(), (nclouds + 1) * 8, hipMemcpyHostToDevice, s));
    TTG_HIP(hipMemcpyAsync(b.d_row_base, b.row_base.data(), (nclouds + 1) * 8, hipMemcpyHostToDevice, s));
    TTG_HIP(hipMemcpyAsync(b.d_tile_base, b.tile_base.data(), (nclouds + 1) * 8, hipMemcpyHostToDevice, s));
    TTG_HIP(hipMemcpyAsync(b.d_pair_base, b.pair_base.data(), (nclouds + 1) * 8, hipMemcpyHostToDevice, s));
    TTG_HIP(hipMemcpyAsync(b.d_bmin_off, b.bmin_off.data(), (nclouds + 1) * 8, hipMemcpyHostToDevice, s));
    {   // per tile and per row, the cloud it belongs to (one load instead of a search)
        std::vector<u32> tc(std::max<u64>(b.total_tiles, 1)), rc(std::max<u64>(b.total_rows, 1));
        for (u32 c = 0; c < nclouds; ++c) {
            std::fill(tc.begin() + b.tile_base[c], tc.begin() + b.tile_base[c] + ntiles[c], c);
            std::fill(rc.begin() + b.row_base[c], rc.begin() + b.row_base[c + 1], c);
        }
        TTG_HIP(hipMemcpyAsync(b.d_tile_cloud, tc.data(), tc.size() * 4, hipMemcpyHostToDevice, s));
        TTG_HIP(hipMemcpyAsync(b.d_row_cloud, rc.data(), rc.size() * 4, hipMemcpyHostToDevice, s));
        TTG_HIP(hipStreamSynchronize(s));
    }
    TTG_HIP(hipStreamSynchronize(s));
    b.shape = BatchShape{nclouds, b.d_n, b.d_dim, b.d_pt_off, b.d_mat_off, b.d_row_base, b.d_tile_base,
                         b.d_pair_base, b.d_bmin_off, b.d_tile_cloud, b.d_row_cloud, b.total_tiles,
                         b.total_rows};
    TTG_HIP(hipHostMalloc(reinterpret_cast<void**>(&b.h_pts), std::max<u64>(b.total_pts, 1) * 4, hipHostMallocDefault));
    TTG_HIP(hipHostMalloc(reinterpret_cast<void**>(&b.h_mats), std::max<u64>(b.total_mat, 1) * 4, hipHostMallocDefault));
    b.ev.resize(nclouds);
    for (auto& e : b.ev) TTG_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    TTG_HIP(hipEventCreate(&b.t0));
    TTG_HIP(hipEventCreate(&b.t1));
    b.seg.resize(nclouds + 1);
    b.kseg.resize(nclouds + 1);
    b.thr.resize(nclouds);
}

// Times each tile shape on each dimension group of this batch (points already
// on the device): one warm-up launch, then the fastest of three, per shape;
// keeps the fastest shape of every group for this (metric variant,
// dimension set). Results never depend on the shapes.
void batch_autotune(BatchState& b) {
    u64 dimkey = 0;
    for (u32 d : b.dim) dimkey |= u64(1) << std::min<u32>(d, 63);
    const auto key = std::make_pair(int(b.variant), dimkey);
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune.find(key);
        if (it != g_tune.end()) {
            std::copy(it->second.rpw, it->second.rpw + kDimGroups, b.rpw_g);
            b.tuned = true;
            return;
        }
    }
    TileShapes ts{};
    fill_device_u32(b.d_rowmax, std::max<u64>(b.total_rows, 1), 0u, b.stream);
    for (int g = 0; g < kDimGroups; ++g) {
        ts.rpw[g] = 16;
        if (b.group_tiles[g + 1] <= b.group_tiles[g]) continue;
        float best = 1e30f;
        for (int rpw : {8, 16, 32}) {
            for (int rep = 0; rep < 4; ++rep) {   // the first launch warms up
                TTG_HIP(hipEventRecord(b.t0, b.stream));
                launch_batch_dist_v(b.variant, g, rpw, nullptr, b.group_tiles, b.stream, b.shape, b.ma, b.d_pts,
                                    b.d_mats, b.d_rowmax);
                TTG_HIP(hipEventRecord(b.t1, b.stream));
                TTG_HIP(hipEventSynchronize(b.t1));
                float ms = 0;
                TTG_HIP(hipEventElapsedTime(&ms, b.t0, b.t1));
                if (rep > 0 && ms < best) {
                    best = ms;
                    ts.rpw[g] = rpw;
                }
            }
        }
    }
    std::copy(ts.rpw, ts.rpw + kDimGroups, b.rpw_g);
    b.tuned = true;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[key] = ts;
}

void batch_run(BatchState& b, const float* pts, int use_graph) {
    hipStream_t s = b.stream;
    if (b.total_pts) {
        std::memcpy(b.h_pts, pts, b.total_pts * 4);
        TTG_HIP(hipMemcpyAsync(b.d_pts, b.h_pts, b.total_pts * 4, hipMemcpyHostToDevice, s));
    }
    if (!b.tuned) batch_autotune(b);
    TTG_HIP(hipEventRecord(b.t0, s));
    if (use_graph) {
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
    // sizes of the data-dependent outputs
    u32 kept = 0;
    TTG_HIP(hipMemcpyAsync(b.kseg.data(), b.d_kseg, (b.nclouds + 1) * 4, hipMemcpyDeviceToHost, s));
    TTG_HIP(hipMemcpyAsync(b.thr.data(), b.d_thr, b.nclouds * 4, hipMemcpyDeviceToHost, s));
    TTG_HIP(hipMemcpyAsync(&kept, (b.comb || b.tagged) ? b.d_total : b.d_kept, 4, hipMemcpyDeviceToHost, s));
    TTG_HIP(hipStreamSynchronize(s));
    TTG_HIP(hipEventElapsedTime(&b.last_ms, b.t0, b.t1));
    b.kseg[b.nclouds] = kept;
    if (!b.h_idx || b.h_cap < kept) {
        if (b.h_idx) (void)hipHostFree(b.h_idx);
        if (b.h_d) (void)hipHostFree(b.h_d);
        if (b.h_ap) (void)hipHostFree(b.h_ap);
        b.h_idx = nullptr;
        b.h_d = nullptr;
        b.h_ap = nullptr;
        const u64 cap = std::max<u64>(kept, 1) + (kept >> 3);
        TTG_HIP(hipHostMalloc(reinterpret_cast<void**>(&b.h_idx), cap * 8, hipHostMallocDefault));
        TTG_HIP(hipHostMalloc(reinterpret_cast<void**>(&b.h_d), cap * 4, hipHostMallocDefault));
        b.h_cap = cap;
    }
    if (b.tagged && !b.h_ap)
        TTG_HIP(hipHostMalloc(reinterpret_cast<void**>(&b.h_ap), b.h_cap, hipHostMallocDefault));
    // per cloud: matrix + its kept edges, then an event the host waits on
    for (u32 c = 0; c < b.nclouds; ++c) {
        const u64 nc = b.n[c];
        if (nc)   // padded device rows -> the reference's n x n layout
            TTG_HIP(hipMemcpy2DAsync(b.h_mats + b.mat_off[c], nc * 4, b.d_mats + b.dmat_off[c],
                                     u64(row_ld(u32(nc))) * 4, nc * 4, nc, hipMemcpyDeviceToHost, s));
        const u64 e0 = b.kseg[c], e1 = b.kseg[c + 1];
        if (e1 > e0) {
            TTG_HIP(hipMemcpyAsync(b.h_idx + e0, b.d_out_idx + e0, (e1 - e0) * 8, hipMemcpyDeviceToHost, s));
            TTG_HIP(hipMemcpyAsync(b.h_d + e0, b.d_out_d + e0, (e1 - e0) * 4, hipMemcpyDeviceToHost, s));
            if (b.tagged)
                TTG_HIP(hipMemcpyAsync(b.h_ap + e0, b.d_out_ap + e0, e1 - e0, hipMemcpyDeviceToHost, s));
        }
        TTG_HIP(hipEventRecord(b.ev[c], s));
    }
}

}  // namespace

extern "C" {

tt_status tt_batch_create(uint32_t nclouds, const uint32_t* n, const uint32_t* dim, int kind, float p,
                          float threshold, tt_batch** out) {
    return guarded([&] {
        if (!out) fail(TT_EINVAL, "tt_batch_create: null handle");
        *out = nullptr;
        auto* b = new tt_batch;
        try {
            batch_setup(b->st, nclouds, n, dim, kind, p, threshold);
        } catch (...) {
            delete b;
            throw;
        }
        *out = b;
    });
}

tt_status tt_batch_run(tt_batch* b, const float* pts, int use_graph) {
    return guarded([&] {
        if (!b) fail(TT_EINVAL, "tt_batch_run: null handle");
        if (b->st.total_pts && !pts) fail(TT_EINVAL, "tt_batch_run: null points");
        batch_run(b->st, pts, use_graph);
    });
}

tt_status tt_batch_wait(tt_batch* b, uint32_t cloud) {
    return guarded([&] {
        if (!b || cloud >= b->st.nclouds) fail(TT_EINVAL, "tt_batch_wait: bad cloud");
        TTG_HIP(hipEventSynchronize(b->st.ev[cloud]));
    });
}

tt_status tt_batch_cloud(tt_batch* b, uint32_t cloud, const float** mat, float* threshold, uint64_t* m,
                         const uint64_t** idx, const float** d) {
    return guarded([&] {
        if (!b || cloud >= b->st.nclouds) fail(TT_EINVAL, "tt_batch_cloud: bad cloud");
        BatchState& s = b->st;
        if (s.comb) fail(TT_EINVAL, "tt_batch_cloud: the batch hands over every edge; use tt_batch_cloud_edges");
        if (mat) *mat = s.h_mats + s.mat_off[cloud];
        if (threshold) *threshold = s.thr[cloud];
        const u64 e0 = s.kseg[cloud], e1 = s.kseg[cloud + 1];
        if (m) *m = e1 - e0;
        if (idx) *idx = s.h_idx + e0;
        if (d) *d = s.h_d + e0;
    });
}

tt_status tt_batch_cloud_apparent(tt_batch* b, uint32_t cloud, const uint8_t** apparent) {
    return guarded([&] {
        if (!b || cloud >= b->st.nclouds) fail(TT_EINVAL, "tt_batch_cloud_apparent: bad cloud");
        BatchState& s = b->st;
        if (!s.tagged) fail(TT_EINVAL, "tt_batch_cloud_apparent: the batch hands over no apparent flags (order 3)");
        if (apparent) *apparent = s.h_ap + s.kseg[cloud];
    });
}

tt_status tt_batch_set_order(tt_batch* b, int order) {
    return guarded([&] {
        if (!b) fail(TT_EINVAL, "tt_batch_set_order: null handle");
        if (order < 0 || order > 3) fail(TT_EINVAL, "tt_batch_set_order: order must be 0, 1, 2 or 3");
        BatchState& s = b->st;
        const bool comb = order == 1 || order == 2, witness = order == 2 && s.witness_fits;
        const bool tagged = order == 3;
        if (tagged && !s.witness_fits)
            fail(TT_ELENGTH, "tt_batch_set_order: edge keys with an apparent flag exceed 64 bits");
        if (s.comb == comb && s.witness == witness && s.tagged == tagged) return;
        if (tagged && !s.d_out_ap) s.d_out_ap = s.ws.get<u8>(s.capacity);
        s.comb = comb;
        s.witness = witness;
        s.tagged = tagged;
        // the recorded sequence belongs to the other order
        if (s.exec) TTG_HIP(hipGraphExecDestroy(s.exec));
        if (s.graph) TTG_HIP(hipGraphDestroy(s.graph));
        s.exec = nullptr;
        s.graph = nullptr;
    });
}

tt_status tt_batch_cloud_edges(tt_batch* b, uint32_t cloud, const float** mat, float* threshold, uint64_t* m,
                               const uint32_t** verts, const float** d, int* witness_bit) {
    return guarded([&] {
        if (!b || cloud >= b->st.nclouds) fail(TT_EINVAL, "tt_batch_cloud_edges: bad cloud");
        BatchState& s = b->st;
        if (!s.comb) fail(TT_EINVAL, "tt_batch_cloud_edges: the batch hands over the coboundary order");
        if (mat) *mat = s.h_mats + s.mat_off[cloud];
        if (threshold) *threshold = s.thr[cloud];
        const u64 e0 = s.kseg[cloud], e1 = s.kseg[cloud + 1];
        if (m) *m = e1 - e0;
        if (verts) *verts = reinterpret_cast<const uint32_t*>(s.h_idx + e0);
        if (d) *d = s.h_d + e0;
        if (witness_bit) *witness_bit = s.witness ? 1 : 0;
    });
}

tt_status tt_batch_info(tt_batch* b, float* device_ms, int* rows_per_wave, uint64_t* edges_total,
                        uint64_t* edges_kept) {
    return guarded([&] {
        if (!b) fail(TT_EINVAL, "tt_batch_info: null handle");
        if (device_ms) *device_ms = b->st.last_ms;
        if (rows_per_wave) *rows_per_wave = b->st.rpw_main();
        if (edges_total || edges_kept) {
            u32 total = 0;
            TTG_HIP(hipMemcpy(&total, b->st.d_total, 4, hipMemcpyDeviceToHost));
            if (edges_total) *edges_total = total;
            if (edges_kept) *edges_kept = b->st.kseg[b->st.nclouds];
        }
    });
}

tt_status tt_batch_time_dist(tt_batch* b, float* ms) {
    return guarded([&] {
        if (!b || !ms) fail(TT_EINVAL, "tt_batch_time_dist: null argument");
        BatchState& st = b->st;
        if (!st.tuned) fail(TT_EINVAL, "tt_batch_time_dist: run the batch first");
        // one warm-up, then the median of five timed launches of the stage
        // (every dimension group, each with its tuned tile shape)
        std::vector<float> t;
        for (int rep = 0; rep < 6; ++rep) {
            fill_device_u32(st.d_rowmax, std::max<u64>(st.total_rows, 1), 0u, st.stream);
            TTG_HIP(hipEventRecord(st.t0, st.stream));
            if (st.total_tiles)
                launch_batch_dist_v(st.variant, -1, 0, st.rpw_g, st.group_tiles, st.stream, st.shape, st.ma, st.d_pts,
                                    st.d_mats, st.d_rowmax);
            TTG_HIP(hipEventRecord(st.t1, st.stream));
            TTG_HIP(hipEventSynchronize(st.t1));
            float ms1 = 0;
            TTG_HIP(hipEventElapsedTime(&ms1, st.t0, st.t1));
            if (rep) t.push_back(ms1);
        }
        std::sort(t.begin(), t.end());
        *ms = t[t.size() / 2];
    });
}

void tt_batch_destroy(tt_batch* b) { delete b; }

}  // extern "C"
