// SPDX-License-Identifier: MIT
// dist.hip — Minkowski distance matrix in the reducer's full or compressed
// lower-triangular layout, with the enclosing radius fused in.
//
// Replaces rips::make_dense (include/turbotops/rips.hpp:391-416) and
// DenseComplex::enclosing_radius (rips.hpp:370-380).
//
// Work decomposition: one 256-thread workgroup (4 wave64s) per 64x64 tile
// of the LOWER triangle of the n x n matrix (tile row bi >= tile column bj),
// so every distance is evaluated once. Lane l of wave w owns column
// j = 64 bj + l and rows i = 64 bi + 16 w + r (r < 16): the column point sits
// in VGPRs, the row points are wave-uniform and come through scalar loads.
// Each wave stores one 256-byte contiguous row segment per instruction. For
// the full layout the tile is staged in LDS and stored a second time,
// transposed, into the upper triangle (rows j, columns i) — the reference's
// mirror pass (rips.hpp:401-403) without re-reading the matrix.
//
// Enclosing radius: the tile's 64 row maxima and 64 column maxima are
// reduced through LDS and folded into rowmax[] with 32-bit atomicMax on the
// float bit patterns (distances are >= 0, where bit order is value order);
// a one-block kernel then takes the minimum. max and min are exact, so the
// result is the reference's bit for bit whatever the order.
//
// Arithmetic per pair is the reference's scalar kernel (metric.hpp:92-207,
// NO_SIMD branches) step by step; the build uses -ffp-contract=off.
#include <algorithm>

#include <atomic>

#include "frontend.hpp"
#include "tiles.hpp"

namespace ttg {

// Rows per wave of the distance tiles (8, 16 or 32; 512, 256 or 128 threads
// per tile); 0 picks by dimension (below). Results never depend on it.
static std::atomic<int> g_tile_rpw{0};

int tile_rows_per_wave() { return g_tile_rpw.load(); }

// The single-cloud default: 8 rows per wave up to d = 4, 16 beyond. Measured
// on n = 20000: d = 3 L2 0.310 ms at 8 against 0.317 at 16 and 0.318 at 32;
// d = 16 L1 0.334-0.340 ms at 16 against 0.347-0.350 at 8 and 0.401-0.407 at 32.
static int rows_per_wave_for(u32 dim) { return dim <= 4 ? 8 : 16; }
void set_tile_rows_per_wave(int rpw) {
    if (rpw != 8 && rpw != 16 && rpw != 32) fail(TT_EINVAL, "tile rows per wave must be 8, 16 or 32");
    g_tile_rpw.store(rpw);
}

namespace {

template <int V, int DIM, int LAYOUT, int RPW>
__global__ __launch_bounds__(64 * (64 / RPW)) void dist_tiles(const float* __restrict__ x, u32 n,
                                                              u32 dim_rt, MetricArgs ma,
                                                              float* __restrict__ mat,
                                                              u32* __restrict__ rowmax) {
    __shared__ TileSmem<RPW> sm;
    u32 bi, bj;
    lower_tile_of(blockIdx.x, bi, bj);
    dist_tile<V, DIM, LAYOUT, RPW>(x, n, dim_rt, ma, bi, bj, mat, rowmax, sm, n);
}

__global__ void fill_u32(u32* __restrict__ p, u64 count, u32 v) {
    for (u64 t = blockIdx.x * u64(blockDim.x) + threadIdx.x; t < count;
         t += u64(gridDim.x) * blockDim.x)
        p[t] = v;
}

// min_i rowmax[i] as a float; +inf for n == 0 (rips.hpp:377-379).
__global__ __launch_bounds__(1024) void min_rowmax(const u32* __restrict__ rowmax, u32 n,
                                                   float* __restrict__ out) {
    __shared__ u32 red[1024 / 64];
    u32 m = 0x7f800000u;   // +inf
    // eight loads in flight per thread (one at a time left this one-block
    // kernel waiting on each)
    for (u32 t0 = threadIdx.x; t0 < n; t0 += 8 * blockDim.x) {
        u32 v[8];
#pragma unroll
        for (u32 u = 0; u < 8; ++u) {
            const u32 t = t0 + u * blockDim.x;
            v[u] = t < n ? rowmax[t] : 0x7f800000u;
        }
#pragma unroll
        for (u32 u = 0; u < 8; ++u) m = min(m, v[u]);
    }
    for (int off = 32; off > 0; off >>= 1) m = min(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        u32 r = red[0];
        for (u32 w = 1; w < blockDim.x / 64; ++w) r = min(r, red[w]);
        *out = __uint_as_float(r);
    }
}

template <int V, int LAYOUT, int RPW>
void launch_dim(u32 blocks, hipStream_t s, const float* x, u32 n, u32 dim, MetricArgs ma,
                float* mat, u32* rowmax) {
    constexpr u32 threads = 64 * (64 / RPW);
#define TTG_DIST_CASE(D) \
    case D: dist_tiles<V, D, LAYOUT, RPW><<<blocks, threads, 0, s>>>(x, n, dim, ma, mat, rowmax); break;
    switch (dim) {
        TTG_DIST_CASE(1)
        TTG_DIST_CASE(2)
        TTG_DIST_CASE(3)
        TTG_DIST_CASE(4)
        TTG_DIST_CASE(8)
        TTG_DIST_CASE(16)
        default: dist_tiles<V, 0, LAYOUT, RPW><<<blocks, threads, 0, s>>>(x, n, dim, ma, mat, rowmax);
    }
#undef TTG_DIST_CASE
}

template <int LAYOUT, int RPW>
void launch_variant(MetricVariant v, u32 blocks, hipStream_t s, const float* x, u32 n, u32 dim,
                    MetricArgs ma, float* mat, u32* rowmax) {
    switch (v) {
        case MV_L1: launch_dim<MV_L1, LAYOUT, RPW>(blocks, s, x, n, dim, ma, mat, rowmax); break;
        case MV_L2: launch_dim<MV_L2, LAYOUT, RPW>(blocks, s, x, n, dim, ma, mat, rowmax); break;
        case MV_LINF: launch_dim<MV_LINF, LAYOUT, RPW>(blocks, s, x, n, dim, ma, mat, rowmax); break;
        case MV_LP_INT: launch_dim<MV_LP_INT, LAYOUT, RPW>(blocks, s, x, n, dim, ma, mat, rowmax); break;
        default: launch_dim<MV_LP_POW, LAYOUT, RPW>(blocks, s, x, n, dim, ma, mat, rowmax); break;
    }
}

template <int LAYOUT>
void launch_rpw(int rpw, MetricVariant v, u32 blocks, hipStream_t s, const float* x, u32 n, u32 dim,
                MetricArgs ma, float* mat, u32* rowmax) {
    switch (rpw) {
        case 8: launch_variant<LAYOUT, 8>(v, blocks, s, x, n, dim, ma, mat, rowmax); break;
        case 32: launch_variant<LAYOUT, 32>(v, blocks, s, x, n, dim, ma, mat, rowmax); break;
        default: launch_variant<LAYOUT, 16>(v, blocks, s, x, n, dim, ma, mat, rowmax); break;
    }
}

}  // namespace

MetricArgs metric_args(int kind, float p) {
    MetricArgs ma;
    ma.p = p;
    ma.inv_p = 1.0f / p;
    ma.ip = (kind == TT_METRIC_LP && p == std::floor(p) && p <= 8.0f) ? u32(p) : 0u;
    ma.half = 0;
    if (kind == TT_METRIC_LP && !ma.ip && p - std::floor(p) == 0.5f && p <= 16.5f) {
        ma.half = 1;
        ma.ip = u32(std::floor(p));
    }
    return ma;
}

void fill_device_u32(u32* p, u64 count, u32 v, hipStream_t s) {
    if (!count) return;
    const u32 blocks = u32(std::min<u64>((count + 255) / 256, 4096));
    fill_u32<<<blocks, 256, 0, s>>>(p, count, v);
    TTG_LAUNCH_CHECK();
}

// Stream-ordered distance matrix (+ row maxima, + enclosing radius).
void dist_matrix(const float* d_x, u32 n, u32 dim, int kind, float p, int layout, float* d_mat,
                 u32* d_rowmax, float* d_radius, hipStream_t s, int rpw_override,
                 hipEvent_t ev_begin, hipEvent_t ev_end) {
    check_metric(kind, p);
    if (layout != TT_LAYOUT_FULL && layout != TT_LAYOUT_LOWER)
        fail(TT_EINVAL, "dist_matrix: unknown layout " + std::to_string(layout));
    if (n && dim == 0) fail(TT_EINVAL, "dist_matrix: dim must be >= 1");
    const u64 cells = layout == TT_LAYOUT_FULL ? u64(n) * n : u64(n) * (n ? n - 1 : 0) / 2;
    if ((n && !d_x) || (cells && !d_mat)) fail(TT_EINVAL, "dist_matrix: null buffer");
    if (d_radius && !d_rowmax) fail(TT_EINVAL, "dist_matrix: the radius needs a rowmax buffer");
    if (d_rowmax) fill_device_u32(d_rowmax, n, 0u, s);
    if (n) {
        const u64 nb = (u64(n) + DTILE - 1) / DTILE;
        const u64 blocks = nb * (nb + 1) / 2;
        if (blocks > 0x7fffffffull) fail(TT_ELENGTH, "dist_matrix: n too large for one launch");
        const MetricVariant v = variant_of(kind, p);
        const MetricArgs ma = metric_args(kind, p);
        const int rpw = rpw_override > 0 ? rpw_override
                        : tile_rows_per_wave() > 0 ? tile_rows_per_wave() : rows_per_wave_for(dim);
        if (ev_begin) TTG_HIP(hipEventRecord(ev_begin, s));
        if (layout == TT_LAYOUT_FULL)
            launch_rpw<TT_LAYOUT_FULL>(rpw, v, u32(blocks), s, d_x, n, dim, ma, d_mat, d_rowmax);
        else
            launch_rpw<TT_LAYOUT_LOWER>(rpw, v, u32(blocks), s, d_x, n, dim, ma, d_mat, d_rowmax);
        TTG_LAUNCH_CHECK();
        if (ev_end) TTG_HIP(hipEventRecord(ev_end, s));
    }
    if (d_radius) {
        min_rowmax<<<1, 1024, 0, s>>>(d_rowmax, n, d_radius);
        TTG_LAUNCH_CHECK();
    }
}

}  // namespace ttg

using namespace ttg;

extern "C" {

tt_status tt_dist_matrix_dev(const float* d_pts, uint32_t n, uint32_t dim, int kind, float p,
                             int layout, float* d_mat, uint32_t* d_rowmax, float* d_radius,
                             void* stream) {
    return guarded([&] {
        dist_matrix(d_pts, n, dim, kind, p, layout, d_mat, d_rowmax, d_radius, as_stream(stream));
    });
}

tt_status tt_dist_matrix_dev_events(const float* d_pts, uint32_t n, uint32_t dim, int kind, float p,
                                    int layout, float* d_mat, uint32_t* d_rowmax, float* d_radius,
                                    void* stream, void* ev_begin, void* ev_end) {
    return guarded([&] {
        if (!n && (ev_begin || ev_end)) fail(TT_EINVAL, "tt_dist_matrix_dev_events: no tile launch to time at n = 0");
        dist_matrix(d_pts, n, dim, kind, p, layout, d_mat, d_rowmax, d_radius, as_stream(stream), 0,
                    static_cast<hipEvent_t>(ev_begin), static_cast<hipEvent_t>(ev_end));
    });
}

tt_status tt_make_dense(const float* pts, uint32_t n, uint32_t dim, int kind, float p, int layout,
                        float* mat, float* radius) {
    return guarded([&] {
        check_metric(kind, p);
        check_points(pts, u64(n) * dim, "tt_make_dense");
        const u64 cells = layout == TT_LAYOUT_FULL ? u64(n) * n : u64(n) * (n ? n - 1 : 0) / 2;
        if (cells && !mat) fail(TT_EINVAL, "tt_make_dense: null matrix");
        DevBuf<float> dx(u64(n) * dim), dm(cells), dr(1);
        DevBuf<u32> rowmax(radius ? n : 0);
        hipStream_t s = nullptr;
        if (dx.n) TTG_HIP(hipMemcpy(dx.p, pts, dx.n * sizeof(float), hipMemcpyHostToDevice));
        dist_matrix(dx.p, n, dim, kind, p, layout, dm.p, radius ? rowmax.p : nullptr,
                    radius ? dr.p : nullptr, s);
        copy_to_host(mat, dm.p, cells * sizeof(float), s);
        if (radius) TTG_HIP(hipMemcpy(radius, dr.p, sizeof(float), hipMemcpyDeviceToHost));
    });
}

}  // extern "C"
