// SPDX-License-Identifier: MIT
// edges.hip — the coboundary reducer's dimension-0 inputs on the dense
// complex: the kept edges in filtration order and, per edge, the apparent
// pair test.
//
// Replaces Reducer::dim0's edge enumeration, counting and sort
// (include/turbotops/rips.hpp:685-839) and its parallel apparent test
// (rips.hpp:844-849 -> zero_apparent_cofacet, :585-665 on the full layout,
// cofacets_full :281-344).
//
// Pipeline (all on one stream, matrix resident in HBM, full layout):
//   edge_count     per row i: #{j < i : d(i, j) <= threshold}, plus the min /
//                  max float bits of the kept diameters (exact integer order
//                  of non-negative floats)
//   scan           row offsets
//   edge_compact   one wave per row, ballot + mbcnt compaction of the kept
//                  edges into 64-bit keys  ((d - dmin) << ib) | (M - 1 - idx),
//                  M = n(n-1)/2: ascending keys are the reference's
//                  `precedes` order (diameter ascending, index descending)
//   block_min      per row, the minimum of every block of 64 columns
//   apparent_scan  on the compacted (row-major) edges, 64 edges per wave:
//                  the lanes first walk the top TOPK candidates w together
//                  (coalesced through the matrix's symmetry); an edge still
//                  open is then searched by the whole wave, the block minima of
//                  rows i and j picking the blocks that can hold a w with
//                  max(d(i,w), d(j,w)) <= d, highest first. The zero-pivot-facet
//                  rule on the triangle decides whether the edge is its
//                  apparent partner (apparent.hpp). The flag is appended below
//                  the key.
//   radix sort     only the bits that vary (plus the flag bit)
#include <algorithm>

#include "apparent.hpp"
#include "frontend.hpp"

namespace ttg {

namespace {

constexpr u32 EB = 256;

// Largest v with v(v-1)/2 <= idx, and the edge (v, idx - v(v-1)/2), as the
// reference's closed-form decode (rips.hpp:549-556).
__device__ __forceinline__ void decode_edge(u64 idx, u32& i, u32& j) {
    u64 v = u64((1.0 + __builtin_sqrt(1.0 + 8.0 * double(idx))) * 0.5);
    while (v * (v - 1) / 2 > idx) --v;
    while ((v + 1) * v / 2 <= idx) ++v;
    i = u32(v);
    j = u32(idx - v * (v - 1) / 2);
}

// Rows are taken longest first (row i holds i kept-edge candidates), so the
// last waves of the grid get the short rows. Each step of a wave issues
// ECH loads of 64 columns before it uses any, so a wave keeps ECH x 256
// bytes in flight instead of one segment (one segment per wave left the
// kernel latency-bound at a fifth of the read rate).
constexpr u32 ECH = 8;

__device__ __forceinline__ u32 row_longest_first(u32 r, u32 n) { return n - 1 - r; }

__global__ __launch_bounds__(EB) void edge_count(const float* __restrict__ mat, u32 n, float thr,
                                                 u64* __restrict__ cnt, u32* __restrict__ mm) {
    // one wave per row, rows grid-strided
    const u32 lane = threadIdx.x & 63;
    const u32 gw = (blockIdx.x * EB + threadIdx.x) >> 6;
    const u32 nw = (gridDim.x * EB) >> 6;
    u32 lo = 0xffffffffu, hi = 0u;
    for (u32 r = gw; r < n; r += nw) {
        const u32 i = row_longest_first(r, n);
        const float* __restrict__ row = mat + u64(i) * n;
        u32 c = 0;
        for (u32 j0 = 0; j0 < i; j0 += 64 * ECH) {
            float d[ECH];
#pragma unroll
            for (u32 u = 0; u < ECH; ++u) {
                const u32 j = j0 + 64 * u + lane;
                d[u] = j < i ? row[j] : 0.0f;
            }
#pragma unroll
            for (u32 u = 0; u < ECH; ++u) {
                if (j0 + 64 * u + lane < i && d[u] <= thr) {
                    ++c;
                    const u32 b = __float_as_uint(d[u]);
                    lo = min(lo, b);
                    hi = max(hi, b);
                }
            }
        }
#pragma unroll
        for (u32 off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
        if (lane == 0) cnt[i] = c;
    }
#pragma unroll
    for (u32 off = 32; off > 0; off >>= 1) {
        lo = min(lo, __shfl_xor(lo, off));
        hi = max(hi, __shfl_xor(hi, off));
    }
    // Every wave folding its extremes into the same two words serialised
    // 4e4 atomics on one cache line (half of the kernel's time at n = 2e4):
    // a wave now skips the atomic unless its value improves on what it reads
    // (a stale read only costs a redundant atomic; min and max are exact).
    if (lane == 0) {
        const u32 cur_lo = __hip_atomic_load(mm, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const u32 cur_hi = __hip_atomic_load(mm + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (lo < cur_lo) atomicMin(mm, lo);
        if (hi > cur_hi) atomicMax(mm + 1, hi);
    }
}

// Keys in row-major order, which is ascending edge index (comb), or mirrored
// over [0, m), descending index (the coboundary order): the sort then runs
// over the diameter bits alone and keeps the index order of equal diameters
// by stability.
__global__ __launch_bounds__(EB) void edge_compact(const float* __restrict__ mat, u32 n, float thr,
                                                   const u64* __restrict__ off, u32 dmin, int ib,
                                                   u64 M, u64 m, bool comb, u64* __restrict__ keys) {
    const u32 lane = threadIdx.x & 63;
    const u32 gw = (blockIdx.x * EB + threadIdx.x) >> 6;
    const u32 nw = (gridDim.x * EB) >> 6;
    for (u32 r = gw; r < n; r += nw) {
        const u32 i = row_longest_first(r, n);
        const float* __restrict__ row = mat + u64(i) * n;
        u64 o = off[i];
        const u64 rowbase = u64(i) * (i - 1) / 2;
        for (u32 j0 = 0; j0 < i; j0 += 64 * ECH) {
            float d[ECH];
#pragma unroll
            for (u32 u = 0; u < ECH; ++u) {
                const u32 j = j0 + 64 * u + lane;
                d[u] = j < i ? row[j] : 0.0f;
            }
            // chunks in column order: the keys of a row stay in ascending j
#pragma unroll
            for (u32 u = 0; u < ECH; ++u) {
                const u32 j = j0 + 64 * u + lane;
                const bool keep = j < i && d[u] <= thr;
                const u64 b = __ballot(keep);
                if (keep) {
                    const u32 rk = __builtin_amdgcn_mbcnt_hi(u32(b >> 32), __builtin_amdgcn_mbcnt_lo(u32(b), 0u));
                    const u64 idx = rowbase + j;
                    const u64 at = comb ? o + rk : m - 1 - (o + rk);
                    keys[at] = (u64(__float_as_uint(d[u]) - dmin) << ib) | (comb ? idx : M - 1 - idx);
                }
                o += __popcll(b);
            }
        }
    }
}

// Sorted keys -> (idx, d), and the apparent / witness flags where asked for:
// from bit 0 of each key (fb = 1), or from the payload sorted alongside
// (`sorted`). One pass over the keys.
__global__ __launch_bounds__(EB) void edge_unpack(const u64* __restrict__ keys, u64 m, u32 dmin,
                                                  int ib, u64 M, bool comb, int fb,
                                                  const u32* __restrict__ sorted, u8* __restrict__ flags,
                                                  u64* __restrict__ idx, float* __restrict__ d) {
    const u64 mask = ib >= 64 ? ~0ull : ((u64(1) << ib) - 1);
    for (u64 t = blockIdx.x * u64(EB) + threadIdx.x; t < m; t += u64(gridDim.x) * EB) {
        u64 k = keys[t];
        if (flags) flags[t] = sorted ? u8(sorted[t]) : u8(k & 1);
        k >>= fb;
        if (idx) idx[t] = comb ? (k & mask) : M - 1 - (k & mask);
        if (d) d[t] = __uint_as_float(u32(k >> ib) + dmin);
    }
}

// EdgeList vertex pairs (smaller id first) of sorted edge indices.
__global__ __launch_bounds__(EB) void edge_verts(const u64* __restrict__ idx, u64 m, const u8* __restrict__ flags,
                                                 int32_t* __restrict__ verts) {
    for (u64 t = blockIdx.x * u64(EB) + threadIdx.x; t < m; t += u64(gridDim.x) * EB) {
        u32 i, j;
        decode_edge(idx[t], i, j);
        verts[2 * t] = int32_t(j);
        verts[2 * t + 1] = int32_t(i | (flags && flags[t] ? 0x80000000u : 0u));
    }
}

// Block minima of every row of the full matrix, one wave per row.
__global__ __launch_bounds__(EB) void block_min(const float* __restrict__ mat, u32 n, u32 nb,
                                                float* __restrict__ bmin) {
    const u32 lane = threadIdx.x & 63;
    const u32 gw = (blockIdx.x * EB + threadIdx.x) >> 6;
    const u32 nw = (gridDim.x * EB) >> 6;
    for (u32 i = gw; i < n; i += nw) block_min_row(mat + u64(i) * n, n, nb, bmin + u64(i) * nb, lane);
}

// Zero-apparent test of the compacted (row-major) edges, 64 edges per wave,
// one per lane (apparent.hpp); the flag becomes the key's lowest bit. The
// test does not depend on the order, and a flag bit below a unique key does
// not change the order of the sort that follows.
__global__ __launch_bounds__(EB) void apparent_scan(const float* __restrict__ mat, u32 n,
                                                    const float* __restrict__ bmin, u32 nb,
                                                    u64* __restrict__ keys, u64 m, int ib, u64 M,
                                                    u32* __restrict__ flags) {
    const u32 lane = threadIdx.x & 63;
    const u64 gw = (blockIdx.x * u64(EB) + threadIdx.x) >> 6;
    const u64 nw = (u64(gridDim.x) * EB) >> 6;
    const u64 mask = (u64(1) << ib) - 1;
    for (u64 e0 = gw * 64; e0 < m; e0 += nw * 64) {
        const u64 e = e0 + lane;
        const bool valid = e < m;
        u64 k = 0;
        u32 i = 1, j = 0;
        float d = 0.0f;
        if (valid) {
            k = keys[e];
            decode_edge(M - 1 - (k & mask), i, j);
            d = mat[u64(i) * n + j];
        }
        const bool ap = edges_apparent_lanes<TOPK>(mat, bmin, n, n, nb, valid, i, j, d, lane);
        if (valid) {
            if (flags) flags[e] = ap ? 1u : 0u;
            else keys[e] = (k << 1) | (ap ? 1u : 0u);
        }
    }
}

// The boundary engine's ready-witness flag of the compacted comb-order edges
// (keys hold the combinatorial index itself), 64 edges per wave; the flag
// becomes the key's lowest bit, as in apparent_scan.
__global__ __launch_bounds__(EB) void witness_scan(const float* __restrict__ mat, u32 n,
                                                   const float* __restrict__ bmin, u32 nb,
                                                   u64* __restrict__ keys, u64 m, int ib,
                                                   u32* __restrict__ flags) {
    const u32 lane = threadIdx.x & 63;
    const u64 gw = (blockIdx.x * u64(EB) + threadIdx.x) >> 6;
    const u64 nw = (u64(gridDim.x) * EB) >> 6;
    const u64 mask = (u64(1) << ib) - 1;
    for (u64 e0 = gw * 64; e0 < m; e0 += nw * 64) {
        const u64 e = e0 + lane;
        const bool valid = e < m;
        u64 k = 0;
        u32 i = 1, j = 0;
        float d = 0.0f;
        if (valid) {
            k = keys[e];
            decode_edge(k & mask, i, j);
            d = mat[u64(i) * n + j];
        }
        const bool wf = edges_apparent_lanes<TOPK, true>(mat, bmin, n, n, nb, valid, i, j, d, lane);
        if (valid) {
            if (flags) flags[e] = wf ? 1u : 0u;
            else keys[e] = (k << 1) | (wf ? 1u : 0u);
        }
    }
}

u32 grid_for(u64 waves) {
    return u32(std::max<u64>(1, std::min<u64>((waves + 3) / 4, 256 * 64)));
}

}  // namespace

// Kept edges of a full device matrix, in filtration order, and optionally
// their apparent flags. Buffers live in ws.
DenseEdges dense_edges(const float* d_mat, u32 n, float thr, bool want_apparent, Workspace& ws,
                       hipStream_t s, bool comb) {
    // want_apparent: in comb order, the ready-witness flags
    DenseEdges out;
    if (n < 2) return out;
    const u64 M = u64(n) * (n - 1) / 2;   // may pass 2^32 (n > 92682): counts and offsets are 64-bit
    u64* cnt = ws.get<u64>(u64(n) + 1);
    u32* mm = ws.get<u32>(2);
    const u32 init[2] = {0xffffffffu, 0u};
    TTG_HIP(hipMemcpyAsync(mm, init, sizeof(init), hipMemcpyHostToDevice, s));
    edge_count<<<grid_for(n), EB, 0, s>>>(d_mat, n, thr, cnt, mm);
    TTG_LAUNCH_CHECK();
    u64* total = ws.get<u64>(1);
    exclusive_scan_u64(cnt, cnt, n, total, ws, s);
    u32 hm[2];
    u64 hm_total = 0;
    TTG_HIP(hipMemcpyAsync(hm, mm, 2 * sizeof(u32), hipMemcpyDeviceToHost, s));
    TTG_HIP(hipMemcpyAsync(&hm_total, total, sizeof(u64), hipMemcpyDeviceToHost, s));
    TTG_HIP(hipStreamSynchronize(s));
    out.m = hm_total;
    if (!out.m) return out;
    const u32 dmin = hm[0], dmax = hm[1];
    const int ib = std::max(1, bit_length(M - 1));
    const int db = bit_length(u64(dmax - dmin));
    if (ib + db > 64) fail(TT_ELENGTH, "dense_edges: key does not fit 64 bits");
    u64* keys = ws.get<u64>(out.m);
    u64* alt = ws.get<u64>(out.m);
    edge_compact<<<grid_for(n), EB, 0, s>>>(d_mat, n, thr, cnt, dmin, ib, M, out.m, comb, keys);
    TTG_LAUNCH_CHECK();
    // The flag rides below the key where a bit is spare; otherwise (large n
    // and a wide diameter range) as a 32-bit payload of the sort.
    const bool payload = want_apparent && ib + db + 1 > 64;
    const int fb = want_apparent && !payload ? 1 : 0;
    u32* fl = payload ? ws.get<u32>(out.m) : nullptr;
    u32* fl_alt = payload ? ws.get<u32>(out.m) : nullptr;
    if (want_apparent) {
        const u32 nb = (n + BLK - 1) / BLK;
        float* bmin = ws.get<float>(u64(n) * nb);
        block_min<<<grid_for(n), EB, 0, s>>>(d_mat, n, nb, bmin);
        TTG_LAUNCH_CHECK();
        const u32 g = grid_for((out.m + 63) / 64);
        if (comb) witness_scan<<<g, EB, 0, s>>>(d_mat, n, bmin, nb, keys, out.m, ib, fl);
        else apparent_scan<<<g, EB, 0, s>>>(d_mat, n, bmin, nb, keys, out.m, ib, M, fl);
        TTG_LAUNCH_CHECK();
    }
    radix_sort_u64(keys, fl, out.m, ib + fb, ib + db + fb, alt, fl_alt, ws, s);
    const u32 ug = u32(std::min<u64>((out.m + EB - 1) / EB, 65536));
    if (want_apparent) out.apparent = ws.get<u8>(out.m);
    out.idx = alt;   // reuse the ping-pong buffer for the unpacked indices
    out.d = ws.get<float>(out.m);
    edge_unpack<<<ug, EB, 0, s>>>(keys, out.m, dmin, ib, M, comb, fb, fl, out.apparent, out.idx, out.d);
    TTG_LAUNCH_CHECK();
    return out;
}

}  // namespace ttg

using namespace ttg;

extern "C" {

tt_status tt_dense_edges_dev(const float* d_mat, uint32_t n, float threshold, int want_apparent,
                             uint64_t capacity, uint64_t* m, uint64_t* d_idx, float* d_d,
                             uint8_t* d_apparent, void* stream) {
    return guarded([&] {
        if (!m) fail(TT_EINVAL, "tt_dense_edges_dev: null count");
        if (n && !d_mat) fail(TT_EINVAL, "tt_dense_edges_dev: null matrix");
        hipStream_t s = as_stream(stream);
        Workspace ws;
        const DenseEdges e = dense_edges(d_mat, n, threshold, want_apparent != 0, ws, s);
        *m = e.m;
        if (!d_idx) {
            TTG_HIP(hipStreamSynchronize(s));
            return;
        }
        if (e.m > capacity) fail(TT_ELENGTH, "tt_dense_edges_dev: output capacity too small");
        if (e.m) {
            TTG_HIP(hipMemcpyAsync(d_idx, e.idx, e.m * sizeof(u64), hipMemcpyDeviceToDevice, s));
            if (d_d) TTG_HIP(hipMemcpyAsync(d_d, e.d, e.m * sizeof(float), hipMemcpyDeviceToDevice, s));
            if (d_apparent && e.apparent)
                TTG_HIP(hipMemcpyAsync(d_apparent, e.apparent, e.m, hipMemcpyDeviceToDevice, s));
        }
        TTG_HIP(hipStreamSynchronize(s));   // ws is released on return
    });
}

tt_status tt_dense_edges(const float* pts, uint32_t n, uint32_t dim, int kind, float p,
                         float threshold, uint64_t* m, uint64_t** idx, float** d,
                         uint8_t** apparent, float* threshold_out) {
    return guarded([&] {
        check_metric(kind, p);
        check_points(pts, u64(n) * dim, "tt_dense_edges");
        if (!m || !idx || !d) fail(TT_EINVAL, "tt_dense_edges: null output");
        if (std::isnan(threshold) || threshold < 0.0f)
            fail(TT_EINVAL, "persistence: threshold must be non-negative or +inf");
        *m = 0;
        *idx = nullptr;
        *d = nullptr;
        if (apparent) *apparent = nullptr;
        hipStream_t s = nullptr;
        Workspace ws;
        float* dx = ws.get<float>(u64(n) * dim);
        float* dm = ws.get<float>(u64(n) * n);
        u32* rowmax = ws.get<u32>(n);
        float* dr = ws.get<float>(1);
        if (u64(n) * dim)
            TTG_HIP(hipMemcpyAsync(dx, pts, u64(n) * dim * sizeof(float), hipMemcpyHostToDevice, s));
        dist_matrix(dx, n, dim, kind, p, TT_LAYOUT_FULL, dm, rowmax, dr, s);
        float thr = threshold;
        if (std::isinf(threshold)) {
            TTG_HIP(hipMemcpyAsync(&thr, dr, sizeof(float), hipMemcpyDeviceToHost, s));
            TTG_HIP(hipStreamSynchronize(s));
        }
        if (threshold_out) *threshold_out = thr;
        const DenseEdges e = dense_edges(dm, n, thr, apparent != nullptr, ws, s);
        *m = e.m;
        *idx = static_cast<uint64_t*>(std::malloc(std::max<u64>(1, e.m) * sizeof(u64)));
        *d = static_cast<float*>(std::malloc(std::max<u64>(1, e.m) * sizeof(float)));
        if (apparent) *apparent = static_cast<uint8_t*>(std::malloc(std::max<u64>(1, e.m)));
        if (!*idx || !*d || (apparent && !*apparent)) fail(TT_EDEVICE, "out of host memory");
        if (e.m) {
            copy_to_host(*idx, e.idx, e.m * sizeof(u64), s);
            copy_to_host(*d, e.d, e.m * sizeof(float), s);
            if (apparent) copy_to_host(*apparent, e.apparent, e.m, s);
        }
        TTG_HIP(hipStreamSynchronize(s));
    });
}

void tt_free(void* p) { std::free(p); }

}  // extern "C"

namespace ttg {

void edge_vertices(const u64* d_idx, u64 m, int32_t* d_verts, hipStream_t s, const u8* d_flags) {
    if (!m) return;
    edge_verts<<<u32(std::min<u64>((m + EB - 1) / EB, 65536)), EB, 0, s>>>(d_idx, m, d_flags, d_verts);
    TTG_LAUNCH_CHECK();
}

}  // namespace ttg

// The whole dense front end of compute_dense / persistence's dense branch
// (rips.hpp:391-416, :370-380, :1073, :685-849) in one device pass: the
// matrix (copied to `mat` in the requested host layout), the enclosing radius
// as threshold when `threshold` is +inf, the kept edges in filtration order
// and their apparent flags.
extern "C" tt_status tt_dense_filtration(const float* pts, uint32_t n, uint32_t dim, int kind,
                                         float p, float threshold, int layout, float* mat,
                                         uint64_t* m, uint64_t** idx, float** d,
                                         uint8_t** apparent, float* threshold_out) {
    using namespace ttg;
    return guarded([&] {
        check_metric(kind, p);
        check_points(pts, u64(n) * dim, "tt_dense_filtration");
        if (!m || !idx || !d) fail(TT_EINVAL, "tt_dense_filtration: null output");
        if (std::isnan(threshold) || threshold < 0.0f)
            fail(TT_EINVAL, "persistence: threshold must be non-negative or +inf");
        if (layout != TT_LAYOUT_FULL && layout != TT_LAYOUT_LOWER)
            fail(TT_EINVAL, "tt_dense_filtration: unknown layout");
        *m = 0;
        *idx = nullptr;
        *d = nullptr;
        if (apparent) *apparent = nullptr;
        hipStream_t s = nullptr;
        Workspace ws;
        float* dx = ws.get<float>(u64(n) * dim);
        float* dm = ws.get<float>(u64(n) * n);
        u32* rowmax = ws.get<u32>(n);
        float* dr = ws.get<float>(1);
        if (u64(n) * dim)
            TTG_HIP(hipMemcpyAsync(dx, pts, u64(n) * dim * sizeof(float), hipMemcpyHostToDevice, s));
        dist_matrix(dx, n, dim, kind, p, TT_LAYOUT_FULL, dm, rowmax, dr, s);
        float thr = threshold;
        if (std::isinf(threshold)) {
            TTG_HIP(hipMemcpyAsync(&thr, dr, sizeof(float), hipMemcpyDeviceToHost, s));
            TTG_HIP(hipStreamSynchronize(s));
        }
        if (threshold_out) *threshold_out = thr;
        const DenseEdges e = dense_edges(dm, n, thr, apparent != nullptr, ws, s);
        if (mat) {
            if (layout == TT_LAYOUT_FULL) {
                if (u64(n) * n)
                    copy_to_host(mat, dm, u64(n) * n * sizeof(float), s);
            } else {
                const u64 cells = u64(n) * (n ? n - 1 : 0) / 2;
                float* low = ws.get<float>(cells);
                dist_matrix(dx, n, dim, kind, p, TT_LAYOUT_LOWER, low, nullptr, nullptr, s);
                if (cells) copy_to_host(mat, low, cells * sizeof(float), s);
            }
        }
        *m = e.m;
        *idx = static_cast<uint64_t*>(std::malloc(std::max<u64>(1, e.m) * sizeof(u64)));
        *d = static_cast<float*>(std::malloc(std::max<u64>(1, e.m) * sizeof(float)));
        if (apparent) *apparent = static_cast<uint8_t*>(std::malloc(std::max<u64>(1, e.m)));
        if (!*idx || !*d || (apparent && !*apparent)) fail(TT_EDEVICE, "out of host memory");
        if (e.m) {
            copy_to_host(*idx, e.idx, e.m * sizeof(u64), s);
            copy_to_host(*d, e.d, e.m * sizeof(float), s);
            if (apparent) copy_to_host(*apparent, e.apparent, e.m, s);
        }
        TTG_HIP(hipStreamSynchronize(s));
    });
}

// The dense front end of persistence()'s boundary (RedZeD) branch
// (turbotops.hpp:243-246: make_dense, enclosing_radius as threshold) and
// RedZeD's build_edges over it (redzed.hpp:544-613 with DenseGeom::bucket_row
// :677-705): edges i < j with d <= threshold in (diameter ascending,
// combinatorial index ascending) order, as EdgeList holds them (:436-442).
namespace {
tt_status build_edges_impl(const float* pts, uint32_t n, uint32_t dim, int kind, float p, float threshold,
                           int layout, float* mat, uint64_t* m, int32_t** verts, float** diam,
                           float* threshold_out, bool witness) {
    using namespace ttg;
    return guarded([&] {
        check_metric(kind, p);
        check_points(pts, u64(n) * dim, "tt_build_edges");
        if (!m || !verts || !diam) fail(TT_EINVAL, "tt_build_edges: null output");
        if (std::isnan(threshold) || threshold < 0.0f)
            fail(TT_EINVAL, "persistence: threshold must be non-negative or +inf");
        if (layout != TT_LAYOUT_FULL && layout != TT_LAYOUT_LOWER)
            fail(TT_EINVAL, "tt_build_edges: unknown layout");
        *m = 0;
        *verts = nullptr;
        *diam = nullptr;
        hipStream_t s = nullptr;
        Workspace ws;
        float* dx = ws.get<float>(u64(n) * dim);
        float* dm = ws.get<float>(u64(n) * n);
        u32* rowmax = ws.get<u32>(n);
        float* dr = ws.get<float>(1);
        if (u64(n) * dim)
            TTG_HIP(hipMemcpyAsync(dx, pts, u64(n) * dim * sizeof(float), hipMemcpyHostToDevice, s));
        dist_matrix(dx, n, dim, kind, p, TT_LAYOUT_FULL, dm, rowmax, dr, s);
        float thr = threshold;
        if (std::isinf(threshold)) {
            TTG_HIP(hipMemcpyAsync(&thr, dr, sizeof(float), hipMemcpyDeviceToHost, s));
            TTG_HIP(hipStreamSynchronize(s));
        }
        if (threshold_out) *threshold_out = thr;
        const DenseEdges e = dense_edges(dm, n, thr, witness, ws, s, true);
        int32_t* vd = ws.get<int32_t>(2 * e.m);
        edge_vertices(e.idx, e.m, vd, s, witness ? e.apparent : nullptr);
        if (mat) {
            if (layout == TT_LAYOUT_FULL) {
                if (u64(n) * n)
                    copy_to_host(mat, dm, u64(n) * n * sizeof(float), s);
            } else {
                const u64 cells = u64(n) * (n ? n - 1 : 0) / 2;
                float* low = ws.get<float>(cells);
                dist_matrix(dx, n, dim, kind, p, TT_LAYOUT_LOWER, low, nullptr, nullptr, s);
                if (cells) copy_to_host(mat, low, cells * sizeof(float), s);
            }
        }
        *m = e.m;
        *verts = static_cast<int32_t*>(std::malloc(std::max<u64>(1, 2 * e.m) * sizeof(int32_t)));
        *diam = static_cast<float*>(std::malloc(std::max<u64>(1, e.m) * sizeof(float)));
        if (!*verts || !*diam) fail(TT_EDEVICE, "out of host memory");
        if (e.m) {
            copy_to_host(*verts, vd, 2 * e.m * sizeof(int32_t), s);
            copy_to_host(*diam, e.d, e.m * sizeof(float), s);
        }
        TTG_HIP(hipStreamSynchronize(s));
    });
}
}  // namespace

extern "C" tt_status tt_build_edges(const float* pts, uint32_t n, uint32_t dim, int kind, float p,
                                    float threshold, int layout, float* mat, uint64_t* m,
                                    int32_t** verts, float** diam, float* threshold_out) {
    return build_edges_impl(pts, n, dim, kind, p, threshold, layout, mat, m, verts, diam, threshold_out, false);
}

extern "C" tt_status tt_build_edges_witness(const float* pts, uint32_t n, uint32_t dim, int kind, float p,
                                            float threshold, int layout, float* mat, uint64_t* m,
                                            int32_t** verts, float** diam, float* threshold_out) {
    return build_edges_impl(pts, n, dim, kind, p, threshold, layout, mat, m, verts, diam, threshold_out, true);
}
