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

#include "pairs.hpp"

namespace ttg {
namespace {

constexpr u32 TILE = 64;

constexpr u32 BLOCK = 256;    // 4 wave64s

// Linear block id -> lower-triangle tile (bi >= bj).
__device__ __forceinline__ void tile_of(u64 b, u32& bi, u32& bj) {
    u32 t = u32((__builtin_sqrt(8.0 * double(b) + 1.0) - 1.0) * 0.5);
    while (u64(t) * (t + 1) / 2 > b) --t;
    while (u64(t + 1) * (t + 2) / 2 <= b) ++t;
    bi = t;
    bj = u32(b - u64(t) * (t + 1) / 2);
}

template <int V, int DIM, int LAYOUT>
__global__ __launch_bounds__(BLOCK) void dist_tiles(const float* __restrict__ x, u32 n,
                                                    u32 dim_rt, MetricArgs ma,
                                                    float* __restrict__ mat,
                                                    u32* __restrict__ rowmax) {
    __shared__ float tile[TILE][TILE + 1];
    __shared__ float part_row[BLOCK / 64][TILE];
    __shared__ float part_col[BLOCK / 64][TILE];

    u32 bi, bj;
    tile_of(blockIdx.x, bi, bj);
    const u32 lane = threadIdx.x & 63;
    const u32 wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const u32 j = bj * TILE + lane;
    const u32 jc = j < n ? j : n - 1;
    const u32 ibase = bi * TILE + wave * ROWS;

    float val[ROWS];
    if constexpr (DIM > 0) rows_static<V, DIM>(x, n, ma, ibase, jc, val);
    else rows_runtime<V>(x, n, dim_rt, ma, ibase, jc, val);

    // Row segments: (i, j) with j in this tile column.
    float cmax = 0.0f;
#pragma unroll
    for (u32 r = 0; r < ROWS; ++r) {
        const u32 i = ibase + r;
        const bool in = i < n && j < n;
        const float v = in ? val[r] : 0.0f;
        tile[wave * ROWS + r][lane] = v;
        cmax = fmax_ref(cmax, v);
        if constexpr (LAYOUT == TT_LAYOUT_FULL) {
            if (in) __builtin_nontemporal_store(v, mat + u64(i) * n + j);
        } else {
            if (in && j < i) __builtin_nontemporal_store(v, mat + u64(i) * (i - 1) / 2 + j);
        }
    }
    part_col[wave][lane] = cmax;
    __syncthreads();

    // Transposed pass: lane = tile row, 16 tile columns per wave. Stores the
    // mirror (full layout, off-diagonal tiles) and gathers row maxima.
    const u32 i = bi * TILE + lane;
    float rmax = 0.0f;
#pragma unroll
    for (u32 c = 0; c < ROWS; ++c) {
        const u32 cl = wave * ROWS + c;
        const float v = tile[lane][cl];
        rmax = fmax_ref(rmax, v);
        if constexpr (LAYOUT == TT_LAYOUT_FULL) {
            const u32 jj = bj * TILE + cl;
            if (bi != bj && i < n && jj < n) __builtin_nontemporal_store(v, mat + u64(jj) * n + i);
        }
    }
    part_row[wave][lane] = rmax;
    __syncthreads();

    if (rowmax) {
        if (wave == 0) {
            float m = part_col[0][lane];
#pragma unroll
            for (u32 w = 1; w < BLOCK / 64; ++w) m = fmax_ref(m, part_col[w][lane]);
            if (j < n) atomicMax(rowmax + j, __float_as_uint(m));
        } else if (wave == 1) {
            float m = part_row[0][lane];
#pragma unroll
            for (u32 w = 1; w < BLOCK / 64; ++w) m = fmax_ref(m, part_row[w][lane]);
            if (i < n) atomicMax(rowmax + i, __float_as_uint(m));
        }
    }
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
    for (u32 t = threadIdx.x; t < n; t += blockDim.x) m = min(m, rowmax[t]);
    for (int off = 32; off > 0; off >>= 1) m = min(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        u32 r = red[0];
        for (u32 w = 1; w < blockDim.x / 64; ++w) r = min(r, red[w]);
        *out = __uint_as_float(r);
    }
}

template <int V, int LAYOUT>
void launch_dim(u32 blocks, hipStream_t s, const float* x, u32 n, u32 dim, MetricArgs ma,
                float* mat, u32* rowmax) {
#define TTG_DIST_CASE(D) \
    case D: dist_tiles<V, D, LAYOUT><<<blocks, BLOCK, 0, s>>>(x, n, dim, ma, mat, rowmax); break;
    switch (dim) {
        TTG_DIST_CASE(1)
        TTG_DIST_CASE(2)
        TTG_DIST_CASE(3)
        TTG_DIST_CASE(4)
        TTG_DIST_CASE(8)
        TTG_DIST_CASE(16)
        default: dist_tiles<V, 0, LAYOUT><<<blocks, BLOCK, 0, s>>>(x, n, dim, ma, mat, rowmax);
    }
#undef TTG_DIST_CASE
}

template <int LAYOUT>
void launch_variant(MetricVariant v, u32 blocks, hipStream_t s, const float* x, u32 n, u32 dim,
                    MetricArgs ma, float* mat, u32* rowmax) {
    switch (v) {
        case MV_L1: launch_dim<MV_L1, LAYOUT>(blocks, s, x, n, dim, ma, mat, rowmax); break;
        case MV_L2: launch_dim<MV_L2, LAYOUT>(blocks, s, x, n, dim, ma, mat, rowmax); break;
        case MV_LINF: launch_dim<MV_LINF, LAYOUT>(blocks, s, x, n, dim, ma, mat, rowmax); break;
        case MV_LP_INT: launch_dim<MV_LP_INT, LAYOUT>(blocks, s, x, n, dim, ma, mat, rowmax); break;
        default: launch_dim<MV_LP_POW, LAYOUT>(blocks, s, x, n, dim, ma, mat, rowmax); break;
    }
}

}  // namespace

MetricArgs metric_args(int kind, float p) {
    MetricArgs ma;
    ma.p = p;
    ma.inv_p = 1.0f / p;
    ma.ip = (kind == TT_METRIC_LP && p == std::floor(p) && p <= 8.0f) ? u32(p) : 0u;
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
                 u32* d_rowmax, float* d_radius, hipStream_t s) {
    check_metric(kind, p);
    if (layout != TT_LAYOUT_FULL && layout != TT_LAYOUT_LOWER)
        fail(TT_EINVAL, "dist_matrix: unknown layout " + std::to_string(layout));
    if (n && dim == 0) fail(TT_EINVAL, "dist_matrix: dim must be >= 1");
    const u64 cells = layout == TT_LAYOUT_FULL ? u64(n) * n : u64(n) * (n ? n - 1 : 0) / 2;
    if ((n && !d_x) || (cells && !d_mat)) fail(TT_EINVAL, "dist_matrix: null buffer");
    if (d_radius && !d_rowmax) fail(TT_EINVAL, "dist_matrix: the radius needs a rowmax buffer");
    if (d_rowmax) fill_device_u32(d_rowmax, n, 0u, s);
    if (n) {
        const u64 nb = (u64(n) + TILE - 1) / TILE;
        const u64 blocks = nb * (nb + 1) / 2;
        if (blocks > 0x7fffffffull) fail(TT_ELENGTH, "dist_matrix: n too large for one launch");
        const MetricVariant v = variant_of(kind, p);
        const MetricArgs ma = metric_args(kind, p);
        if (layout == TT_LAYOUT_FULL)
            launch_variant<TT_LAYOUT_FULL>(v, u32(blocks), s, d_x, n, dim, ma, d_mat, d_rowmax);
        else
            launch_variant<TT_LAYOUT_LOWER>(v, u32(blocks), s, d_x, n, dim, ma, d_mat, d_rowmax);
        TTG_LAUNCH_CHECK();
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
        if (cells) TTG_HIP(hipMemcpy(mat, dm.p, cells * sizeof(float), hipMemcpyDeviceToHost));
        if (radius) TTG_HIP(hipMemcpy(radius, dr.p, sizeof(float), hipMemcpyDeviceToHost));
    });
}

}  // extern "C"
