// SPDX-License-Identifier: MIT
// sparse.hip — the coboundary reducer's dimension-0 inputs on a sparse
// complex (the flag complex of a weighted graph in CSR form): the kept edges
// in filtration order and, per edge, the apparent pair test.
//
// Replaces Reducer::dim0's edge enumeration and sort for a SparseComplex
// (include/turbotops/rips.hpp:685-839) and its parallel apparent test
// (rips.hpp:844-849 -> zero_apparent_cofacet, :585-665) over
// SparseComplex::cofacets (rips.hpp:428-471: the common neighbours of the
// simplex's vertices, walked from the largest id down).
//
// Pipeline (one stream):
//   sp_count      per row v: #{t : col[t] < v, w[t] <= threshold}
//   scan          row offsets = each edge's position in ascending index order
//                 (rows ascending, columns ascending within a row, so
//                 idx = v(v-1)/2 + u ascends with the position)
//   sp_emit       key = (float bits of w) << 32 | (2^32 - 1 - position): the
//                 ascending keys are the reference's `precedes` order
//                 (diameter ascending, index descending), for any n
//                 (with comb: position ascending, RedZeD's order: diameter
//                 ascending, combinatorial index ascending)
//   radix sort    63 bits
//   sp_unpack     sorted (idx, d)
//   sp_apparent   per edge (a > b, d): the largest common neighbour x with
//                 max(w(a,x), w(b,x)) <= d is the zero-pivot cofacet; the edge
//                 is apparent iff it is that triangle's first facet (ascending
//                 index) of diameter d: x > a -> yes; a > x > b -> iff
//                 w(b,x) != d; x < b -> iff w(a,x) != d and w(b,x) != d.
#include <algorithm>
#include <cmath>
#include <cstdlib>

#include "frontend.hpp"

namespace ttg {

namespace {

constexpr u32 SB = 256;

__global__ void sp_count(const u64* __restrict__ rp, const u32* __restrict__ col, const float* __restrict__ w,
                         u32 n, float thr, u32* __restrict__ cnt) {
    for (u32 v = blockIdx.x * SB + threadIdx.x; v < n; v += gridDim.x * SB) {
        u32 c = 0;
        for (u64 t = rp[v]; t < rp[v + 1]; ++t) c += (col[t] < v) & (w[t] <= thr);
        cnt[v] = c;
    }
}

__global__ void sp_emit(const u64* __restrict__ rp, const u32* __restrict__ col, const float* __restrict__ w,
                        u32 n, float thr, const u32* __restrict__ off, bool comb, u64* __restrict__ keys,
                        u64* __restrict__ pos_idx) {
    for (u32 v = blockIdx.x * SB + threadIdx.x; v < n; v += gridDim.x * SB) {
        u32 o = off[v];
        const u64 base = u64(v) * (v - 1) / 2;
        for (u64 t = rp[v]; t < rp[v + 1]; ++t) {
            const u32 u = col[t];
            const float d = w[t];
            if (u < v && d <= thr) {
                // -0.0 sorts and reports as +0.0 (they compare equal)
                keys[o] = (u64(__float_as_uint(d) & 0x7fffffffu) << 32) | u64(comb ? o : 0xffffffffu - o);
                pos_idx[o] = base + u;
                ++o;
            }
        }
    }
}

__global__ void sp_unpack(const u64* __restrict__ keys, const u64* __restrict__ pos_idx, u64 m, bool comb,
                          u64* __restrict__ idx, float* __restrict__ d) {
    for (u64 e = blockIdx.x * u64(SB) + threadIdx.x; e < m; e += u64(gridDim.x) * SB) {
        const u64 k = keys[e];
        idx[e] = pos_idx[comb ? u32(k) : 0xffffffffu - u32(k)];
        d[e] = __uint_as_float(u32(k >> 32));
    }
}

__device__ __forceinline__ void sp_decode(u64 idx, u32& a, u32& b) {
    u64 v = u64((1.0 + __builtin_sqrt(1.0 + 8.0 * double(idx))) * 0.5);
    while (v * (v - 1) / 2 > idx) --v;
    while ((v + 1) * v / 2 <= idx) ++v;
    a = u32(v);
    b = u32(idx - v * (v - 1) / 2);
}

__global__ void sp_apparent(const u64* __restrict__ rp, const u32* __restrict__ col, const float* __restrict__ w,
                            const u64* __restrict__ idx, const float* __restrict__ dd, u64 m,
                            u8* __restrict__ apparent) {
    for (u64 e = blockIdx.x * u64(SB) + threadIdx.x; e < m; e += u64(gridDim.x) * SB) {
        u32 a, b;
        sp_decode(idx[e], a, b);
        const float d = dd[e];
        // walk both rows from their largest neighbour down to the first
        // common x whose two edges are no longer than d
        u64 pa = rp[a + 1], pb = rp[b + 1];
        const u64 la = rp[a], lb = rp[b];
        u8 res = 0;
        while (pa > la && pb > lb) {
            const u32 xa = col[pa - 1], xb = col[pb - 1];
            if (xa > xb) {
                --pa;
            } else if (xb > xa) {
                --pb;
            } else {
                const float wa = w[pa - 1], wb = w[pb - 1];
                if (wa <= d && wb <= d) {
                    const u32 x = xa;
                    res = x > a ? u8(1) : (x > b ? u8(wb != d) : u8(wb != d && wa != d));
                    break;
                }
                --pa;
                --pb;
            }
        }
        apparent[e] = res;
    }
}

u32 grid_for(u64 items) { return u32(std::max<u64>(1, std::min<u64>((items + SB - 1) / SB, 8192))); }

}  // namespace

SparseEdges sparse_edges(const u64* d_rp, const u32* d_col, const float* d_w, u32 n, float thr,
                         bool want_apparent, Workspace& ws, hipStream_t s, bool comb) {
    SparseEdges out;
    if (n < 2) return out;
    u32* cnt = ws.get<u32>(u64(n) + 1);
    u32* total = ws.get<u32>(1);
    fill_device_u32(cnt, u64(n) + 1, 0u, s);
    sp_count<<<grid_for(n), SB, 0, s>>>(d_rp, d_col, d_w, n, thr, cnt);
    TTG_LAUNCH_CHECK();
    exclusive_scan_u32(cnt, cnt, u64(n) + 1, total, ws, s);
    u32 m = 0;
    TTG_HIP(hipMemcpyAsync(&m, total, 4, hipMemcpyDeviceToHost, s));
    TTG_HIP(hipStreamSynchronize(s));
    out.m = m;
    if (!m) return out;
    u64* keys = ws.get<u64>(m);
    u64* alt = ws.get<u64>(m);
    u64* pos_idx = ws.get<u64>(m);
    sp_emit<<<grid_for(n), SB, 0, s>>>(d_rp, d_col, d_w, n, thr, cnt, comb, keys, pos_idx);
    TTG_LAUNCH_CHECK();
    radix_sort_u64(keys, nullptr, m, 0, 63, alt, nullptr, ws, s);
    out.idx = ws.get<u64>(m);
    out.d = ws.get<float>(m);
    sp_unpack<<<grid_for(m), SB, 0, s>>>(keys, pos_idx, m, comb, out.idx, out.d);
    TTG_LAUNCH_CHECK();
    if (want_apparent) {
        out.apparent = ws.get<u8>(m);
        sp_apparent<<<grid_for(m), SB, 0, s>>>(d_rp, d_col, d_w, out.idx, out.d, m, out.apparent);
        TTG_LAUNCH_CHECK();
    }
    return out;
}

}  // namespace ttg

using namespace ttg;

namespace {

// The CSR invariants the kernels rely on (rows in bounds, neighbour ids
// ascending, < n and not the row itself, as CsrGraph::valid requires,
// core.hpp:137-148); checked on the host before anything is uploaded.
void check_csr(u32 n, const uint64_t* rp, const uint32_t* col, const float* w) {
    if (!rp) fail(TT_EINVAL, "sparse: null row_ptr");
    if (rp[0] != 0) fail(TT_EINVAL, "sparse: row_ptr[0] must be 0");
    for (u32 v = 0; v < n; ++v)
        if (rp[v + 1] < rp[v]) fail(TT_EINVAL, "sparse: row_ptr must be non-decreasing");
    if (rp[n] >= (u64(1) << 32)) fail(TT_ELENGTH, "sparse: more than 2^32 adjacency entries");
    if (rp[n] && (!col || !w)) fail(TT_EINVAL, "sparse: null columns");
    for (u32 v = 0; v < n; ++v)
        for (u64 t = rp[v]; t < rp[v + 1]; ++t) {
            if (col[t] >= n) fail(TT_EINVAL, "sparse: column index out of range");
            if (col[t] == v) fail(TT_EINVAL, "sparse: self loop");
            if (t > rp[v] && col[t] <= col[t - 1]) fail(TT_EINVAL, "sparse: rows must be sorted and unique");
            if (!(w[t] >= 0.0f)) fail(TT_EINVAL, "sparse: weights must be non-negative");
        }
}

}  // namespace

extern "C" {

tt_status tt_sparse_edges(uint32_t n, const uint64_t* row_ptr, const uint32_t* col, const float* w,
                          float threshold, uint64_t* m, uint64_t** idx, float** d, uint8_t** apparent) {
    return guarded([&] {
        if (!m || !idx || !d) fail(TT_EINVAL, "tt_sparse_edges: null output");
        if (std::isnan(threshold) || threshold < 0.0f)
            fail(TT_EINVAL, "persistence: threshold must be non-negative or +inf");
        *m = 0;
        *idx = nullptr;
        *d = nullptr;
        if (apparent) *apparent = nullptr;
        check_csr(n, row_ptr, col, w);
        const u64 nnz = row_ptr[n];
        hipStream_t s = nullptr;
        Workspace ws;
        u64* rp_d = ws.get<u64>(u64(n) + 1);
        u32* col_d = ws.get<u32>(nnz);
        float* w_d = ws.get<float>(nnz);
        TTG_HIP(hipMemcpyAsync(rp_d, row_ptr, (u64(n) + 1) * 8, hipMemcpyHostToDevice, s));
        if (nnz) {
            TTG_HIP(hipMemcpyAsync(col_d, col, nnz * 4, hipMemcpyHostToDevice, s));
            TTG_HIP(hipMemcpyAsync(w_d, w, nnz * 4, hipMemcpyHostToDevice, s));
        }
        const SparseEdges e = sparse_edges(rp_d, col_d, w_d, n, threshold, apparent != nullptr, ws, s);
        *m = e.m;
        *idx = static_cast<uint64_t*>(std::malloc(std::max<u64>(1, e.m) * sizeof(u64)));
        *d = static_cast<float*>(std::malloc(std::max<u64>(1, e.m) * sizeof(float)));
        if (apparent) *apparent = static_cast<uint8_t*>(std::malloc(std::max<u64>(1, e.m)));
        if (!*idx || !*d || (apparent && !*apparent)) fail(TT_EDEVICE, "out of host memory");
        if (e.m) {
            TTG_HIP(hipMemcpyAsync(*idx, e.idx, e.m * sizeof(u64), hipMemcpyDeviceToHost, s));
            TTG_HIP(hipMemcpyAsync(*d, e.d, e.m * sizeof(float), hipMemcpyDeviceToHost, s));
            if (apparent) TTG_HIP(hipMemcpyAsync(*apparent, e.apparent, e.m, hipMemcpyDeviceToHost, s));
        }
        TTG_HIP(hipStreamSynchronize(s));
    });
}

// RedZeD's build_edges over a graph (redzed.hpp:544-613 with
// SparseGeom::bucket_row): edges u < v with w <= threshold in (diameter
// ascending, combinatorial index ascending) order, as EdgeList holds them.
tt_status tt_build_edges_csr(uint32_t n, const uint64_t* row_ptr, const uint32_t* col, const float* w,
                             float threshold, uint64_t* m, int32_t** verts, float** diam) {
    return guarded([&] {
        if (!m || !verts || !diam) fail(TT_EINVAL, "tt_build_edges_csr: null output");
        if (std::isnan(threshold) || threshold < 0.0f)
            fail(TT_EINVAL, "persistence: threshold must be non-negative or +inf");
        *m = 0;
        *verts = nullptr;
        *diam = nullptr;
        check_csr(n, row_ptr, col, w);
        const u64 nnz = row_ptr[n];
        hipStream_t s = nullptr;
        Workspace ws;
        u64* rp_d = ws.get<u64>(u64(n) + 1);
        u32* col_d = ws.get<u32>(nnz);
        float* w_d = ws.get<float>(nnz);
        TTG_HIP(hipMemcpyAsync(rp_d, row_ptr, (u64(n) + 1) * 8, hipMemcpyHostToDevice, s));
        if (nnz) {
            TTG_HIP(hipMemcpyAsync(col_d, col, nnz * 4, hipMemcpyHostToDevice, s));
            TTG_HIP(hipMemcpyAsync(w_d, w, nnz * 4, hipMemcpyHostToDevice, s));
        }
        const SparseEdges e = sparse_edges(rp_d, col_d, w_d, n, threshold, false, ws, s, true);
        int32_t* vd = ws.get<int32_t>(2 * e.m);
        edge_vertices(e.idx, e.m, vd, s);
        *m = e.m;
        *verts = static_cast<int32_t*>(std::malloc(std::max<u64>(1, 2 * e.m) * sizeof(int32_t)));
        *diam = static_cast<float*>(std::malloc(std::max<u64>(1, e.m) * sizeof(float)));
        if (!*verts || !*diam) fail(TT_EDEVICE, "out of host memory");
        if (e.m) {
            TTG_HIP(hipMemcpyAsync(*verts, vd, 2 * e.m * sizeof(int32_t), hipMemcpyDeviceToHost, s));
            TTG_HIP(hipMemcpyAsync(*diam, e.d, e.m * sizeof(float), hipMemcpyDeviceToHost, s));
        }
        TTG_HIP(hipStreamSynchronize(s));
    });
}

}  // extern "C"
