// SPDX-License-Identifier: MIT
// tiles.hpp — one 64 x 64 tile of a distance matrix, shared by the
// single-cloud kernel (dist.hip) and the batched persistent kernel
// (batch.hip).
//
// A workgroup of 64/RPW wave64s covers the tile: lane l owns column
// j = 64 bj + l, wave w owns rows i = 64 bi + RPW w + r (r < RPW). The column
// point sits in VGPRs; row points are wave-uniform (scalar loads). Each wave
// stores one 256-byte row segment per instruction. The tile is staged in LDS
// and, for the full layout, stored again transposed into the mirrored block
// (rows j, columns i) — the reference's mirror pass (rips.hpp:401-403)
// without re-reading the matrix. Row and column maxima are reduced through
// LDS and folded into rowmax[] with atomicMax on the float bits (distances
// are >= 0, where bit order is value order).
#pragma once

#include "pairs.hpp"

namespace ttg {

constexpr u32 DTILE = 64;

// Linear id -> lower-triangle tile (bi >= bj).
__device__ __forceinline__ void lower_tile_of(u64 b, u32& bi, u32& bj) {
    u32 t = u32((__builtin_sqrt(8.0 * double(b) + 1.0) - 1.0) * 0.5);
    while (u64(t) * (t + 1) / 2 > b) --t;
    while (u64(t + 1) * (t + 2) / 2 <= b) ++t;
    bi = t;
    bj = u32(b - u64(t) * (t + 1) / 2);
}

template <int RPW>
struct TileSmem {
    static_assert(RPW == 8 || RPW == 16 || RPW == 32, "2 to 8 waves per tile");
    static constexpr u32 NW = DTILE / RPW;   // waves per workgroup
    float tile[DTILE][DTILE + 1];
    float part_row[NW][DTILE];
    float part_col[NW][DTILE];
};

// One 64-float row segment at a wave-uniform base, stored through a buffer
// resource whose range is `valid` floats: lanes at or past it store nothing
// (the hardware's range check is the column guard). The scalar base and a
// 32-bit lane offset replace a 64-bit address add per store; aux 2 is the
// non-temporal policy (gfx94x/gfx950: bit 0 sc0, bit 1 nt, bit 4 sc1), as
// __builtin_nontemporal_store.
__device__ __forceinline__ void row_store(float* base, u32 valid, u32 lane, float v) {
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(base, 0, int(valid * 4u), 0x00020000);
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), rs, int(lane * 4u), 0, 2);
}

// The tile's values to the matrix (and, full layout, to the mirrored block),
// with each wave's column maxima in part_col and row maxima in part_row.
template <int LAYOUT, int RPW, bool WHOLE>
__device__ __forceinline__ void tile_out(const float (&val)[RPW], u32 n, u32 bi, u32 bj,
                                         float* __restrict__ mat, TileSmem<RPW>& sm, u32 ld) {
    const u32 lane = threadIdx.x & 63;
    const u32 wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const u32 j = bj * DTILE + lane;
    const u32 ibase = bi * DTILE + wave * RPW;
    float cmax = 0.0f;
#pragma unroll
    for (u32 r = 0; r < RPW; ++r) {
        const u32 i = ibase + r;
        const bool in = WHOLE || (i < n && j < n);
        const float v = in ? val[r] : 0.0f;
        sm.tile[wave * RPW + r][lane] = v;
        cmax = fmax_acc(cmax, v);
        if constexpr (LAYOUT == TT_LAYOUT_FULL) {
            const u32 valid = WHOLE ? DTILE : (i < n ? min(DTILE, n - bj * DTILE) : 0u);
            row_store(mat + u64(i) * ld + bj * DTILE, valid, lane, v);
        } else {   // columns j < i (bj <= bi: i - 64 bj >= 0)
            const u32 valid = (WHOLE || i < n) ? min(DTILE, i - bj * DTILE) : 0u;
            row_store(mat + u64(i) * (i - 1) / 2 + bj * DTILE, valid, lane, v);
        }
    }
    sm.part_col[wave][lane] = cmax;
    __syncthreads();

    float rmax = 0.0f;
#pragma unroll
    for (u32 c = 0; c < RPW; ++c) {
        const u32 cl = wave * RPW + c;
        const float v = sm.tile[lane][cl];
        rmax = fmax_acc(rmax, v);
        if constexpr (LAYOUT == TT_LAYOUT_FULL) {
            const u32 jj = bj * DTILE + cl;
            const u32 valid = (bi != bj && (WHOLE || jj < n)) ? (WHOLE ? DTILE : min(DTILE, n - bi * DTILE)) : 0u;
            row_store(mat + u64(jj) * ld + bi * DTILE, valid, lane, v);
        }
    }
    sm.part_row[wave][lane] = rmax;
    __syncthreads();
}

// The reference's value for a dense-matrix cell whose L2 sum is NaN: it
// computes d(x_a, x_b) with a the larger index (rips.hpp:397-399) on x86,
// where a - b returns a NaN operand quieted (a first, sign kept) and inf - inf
// the default NaN 0xffc00000, fma(d, d, s) returns d's NaN before s's, and
// the square root passes its NaN on: the cell holds the NaN of the last NaN
// term. The GPU's arithmetic passes NaNs on by other rules (its subtract
// flips the sign of a NaN in b), so NaN results are rebuilt here (the sum is
// NaN exactly when some term is).
__device__ __noinline__ float l2_nan_ref(const float* __restrict__ a, const float* __restrict__ b, u32 dim) {
    u32 r = 0xffc00000u;
    for (u32 k = 0; k < dim; ++k) {
        const u32 ab = __float_as_uint(a[k]), bb = __float_as_uint(b[k]);
        if ((ab & 0x7fffffffu) > 0x7f800000u) r = ab | 0x00400000u;
        else if ((bb & 0x7fffffffu) > 0x7f800000u) r = bb | 0x00400000u;
        else if ((ab & 0x7fffffffu) == 0x7f800000u && ab == bb) r = 0xffc00000u;
    }
    return __uint_as_float(r);
}

// Cells as the reference holds them: 0 on the diagonal, which it never
// computes (a point with a NaN or infinite coordinate gives NaN against
// itself), and for L2 the reference's NaN. Only diagonal tiles test the
// diagonal, and the NaN repair runs when some value of the wave is NaN.
template <int V, int RPW>
__device__ __forceinline__ void dense_cells(float (&val)[RPW], u32 ibase, u32 jc, u32 n, bool diag,
                                            const float* __restrict__ x, u32 dim) {
    if (diag) {
#pragma unroll
        for (u32 r = 0; r < RPW; ++r)
            if (min(ibase + r, n - 1) == jc) val[r] = 0.0f;
    }
    if constexpr (V == MV_L2) {
        bool nan = false;
#pragma unroll
        for (u32 r = 0; r < RPW; ++r) nan |= val[r] != val[r];
        if (nan) {
#pragma unroll
            for (u32 r = 0; r < RPW; ++r) {
                const u32 i = min(ibase + r, n - 1);
                if (val[r] != val[r]) val[r] = l2_nan_ref(x + u64(max(i, jc)) * dim, x + u64(min(i, jc)) * dim, dim);
            }
        }
    }
}

// ld: row stride of the full layout in floats (n, or n rounded up for
// 128-byte aligned rows); unused for the lower layout.
template <int V, int DIM, int LAYOUT, int RPW>
__device__ __forceinline__ void dist_tile(const float* __restrict__ x, u32 n, u32 dim_rt,
                                          const MetricArgs& ma, u32 bi, u32 bj,
                                          float* __restrict__ mat, u32* __restrict__ rowmax,
                                          TileSmem<RPW>& sm, u32 ld) {
    constexpr u32 NW = TileSmem<RPW>::NW;
    const u32 lane = threadIdx.x & 63;
    const u32 wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const u32 j = bj * DTILE + lane;
    const u32 jc = j < n ? j : n - 1;
    const u32 ibase = bi * DTILE + wave * RPW;

    float val[RPW];
    if constexpr (DIM > 0) rows_static<V, DIM, RPW>(x, n, ma, ibase, jc, val);
    else rows_runtime<V, RPW>(x, n, dim_rt, ma, ibase, jc, val);
    dense_cells<V, RPW>(val, ibase, jc, n, bi == bj, x, DIM > 0 ? u32(DIM) : dim_rt);

    // A tile with every row and column below n (bj <= bi, so (bi + 1) * 64 <= n
    // suffices) stores without per-element guards; each store addresses a
    // wave-uniform row base plus the lane, one SGPR pair and a 32-bit offset.
    // The running maxima take fmax_acc: the values are distances (>= +0 or
    // NaN) and 0 for cells outside the matrix.
    if ((bi + 1) * DTILE <= n) tile_out<LAYOUT, RPW, true>(val, n, bi, bj, mat, sm, ld);
    else tile_out<LAYOUT, RPW, false>(val, n, bi, bj, mat, sm, ld);

    const u32 i = bi * DTILE + lane;
    if (rowmax) {
        if (wave == 0) {
            float m = sm.part_col[0][lane];
#pragma unroll
            for (u32 w = 1; w < NW; ++w) m = fmax_ref(m, sm.part_col[w][lane]);
            if (j < n) atomicMax(rowmax + j, __float_as_uint(m));
        } else if (wave == 1) {
            float m = sm.part_row[0][lane];
#pragma unroll
            for (u32 w = 1; w < NW; ++w) m = fmax_ref(m, sm.part_row[w][lane]);
            if (i < n) atomicMax(rowmax + i, __float_as_uint(m));
        }
    }
}

}  // namespace ttg
