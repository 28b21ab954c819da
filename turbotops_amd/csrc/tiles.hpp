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

    float cmax = 0.0f;
#pragma unroll
    for (u32 r = 0; r < RPW; ++r) {
        const u32 i = ibase + r;
        const bool in = i < n && j < n;
        const float v = in ? val[r] : 0.0f;
        sm.tile[wave * RPW + r][lane] = v;
        cmax = fmax_ref(cmax, v);
        if constexpr (LAYOUT == TT_LAYOUT_FULL) {
            if (in) __builtin_nontemporal_store(v, mat + u64(i) * ld + j);
        } else {
            if (in && j < i) __builtin_nontemporal_store(v, mat + u64(i) * (i - 1) / 2 + j);
        }
    }
    sm.part_col[wave][lane] = cmax;
    __syncthreads();

    const u32 i = bi * DTILE + lane;
    float rmax = 0.0f;
#pragma unroll
    for (u32 c = 0; c < RPW; ++c) {
        const u32 cl = wave * RPW + c;
        const float v = sm.tile[lane][cl];
        rmax = fmax_ref(rmax, v);
        if constexpr (LAYOUT == TT_LAYOUT_FULL) {
            const u32 jj = bj * DTILE + cl;
            if (bi != bj && i < n && jj < n) __builtin_nontemporal_store(v, mat + u64(jj) * ld + i);
        }
    }
    sm.part_row[wave][lane] = rmax;
    __syncthreads();

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
