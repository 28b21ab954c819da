// SPDX-License-Identifier: MIT
// apparent.hpp — the zero-apparent test of an edge on a full distance
// matrix, one wave64 per edge, with a block-minimum filter.
//
// Reference semantics (include/turbotops/rips.hpp:585-665 with cofacets_full
// :326-343 and cap = the edge's diameter d): the edge (i, j), i > j, is
// apparent iff its zero-pivot cofacet — the triangle {i, j, w} with the
// largest w != i, j such that max(d(i, w), d(j, w)) <= d — exists and the
// edge is that triangle's first facet of diameter d in ascending index order.
//
// The search walks w downward. A plain walk reads two matrix rows until it
// finds w, up to 2n floats per edge for short edges. Here every row also has
// its minima over blocks of 64 columns, bmin[i][b] = min over w in block b of
// d(i, w); a block holds a candidate only if bmin[i][b] <= d and
// bmin[j][b] <= d, so one 64-lane pass over the block minima picks the
// blocks worth reading, highest first, and only those are read. The filter is
// conservative (a block's minimum includes w = i and w = j), so the first w
// found is exactly the plain walk's.
#pragma once

#include "common.hpp"

namespace ttg {

constexpr u32 BLK = 64;
constexpr u32 TOPK = 64;   // candidates the lanes walk together before the wave search

__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
    for (u32 off = 32; off > 0; off >>= 1) v = fminf(v, __shfl_xor(v, off));
    return v;
}

// bmin_row[b] for every block b of one row (whole wave; lane = lane id).
__device__ __forceinline__ void block_min_row(const float* __restrict__ row, u32 n, u32 nb,
                                              float* __restrict__ bmin_row, u32 lane) {
    for (u32 b = 0; b < nb; ++b) {
        const u32 w = b * BLK + lane;
        const float v = wave_min(w < n ? row[w] : INFINITY);
        if (lane == 0) bmin_row[b] = v;
    }
}

// The zero-pivot-facet rule (rips.hpp:609-636) on the triangle {i, j, w}
// found as the edge's zero-pivot cofacet: the first facet in ascending index
// order holding an edge of length d must be (i, j) itself.
__device__ __forceinline__ bool facet_rule(u32 w, u32 i, u32 j, float d, float diw, float djw) {
    if (w > i) return true;                     // facet (i, j) comes first
    if (w > j) return djw != d;                 // (w, j) precedes (i, j)
    return djw != d && diw != d;                // (j, w), (i, w) precede
}

// Apparent flag of edge (i, j) with diameter d, searching only w < w_hi
// (whole wave; result valid in every lane). ri, rj: rows i and j; bi, bj:
// their block minima.
__device__ __forceinline__ bool edge_is_apparent(const float* __restrict__ ri, const float* __restrict__ rj,
                                                 const float* __restrict__ bi, const float* __restrict__ bj,
                                                 u32 n, u32 nb, u32 i, u32 j, float d, u32 lane,
                                                 u32 w_hi) {
    int wf = -1;
    float diw = 0.0f, djw = 0.0f;
    const int top_blk = w_hi ? int((w_hi - 1) / BLK) : -1;
    (void)nb;
    for (int top = top_blk; top >= 0 && wf < 0; top -= 64) {
        const int b = top - int(lane);
        bool cand = false;
        if (b >= 0) cand = fmax_ref(bi[b], bj[b]) <= d;
        u64 bal = __ballot(cand);
        while (bal) {
            const int blk = top - __builtin_ctzll(bal);   // highest remaining block first
            bal &= bal - 1;
            const int w = blk * int(BLK) + int(BLK - 1 - lane);
            float a = INFINITY, c = INFINITY;
            const bool in = u32(w) < w_hi;              // w_hi <= n
            if (in) {
                a = ri[w];
                c = rj[w];
            }
            const bool ok = in && u32(w) != i && u32(w) != j && fmax_ref(a, c) <= d;
            const u64 hit = __ballot(ok);
            if (hit) {
                const int src = __builtin_ctzll(hit);       // lowest lane = largest w
                wf = blk * int(BLK) + int(BLK - 1) - src;
                diw = __shfl(a, src);
                djw = __shfl(c, src);
                break;
            }
        }
    }
    return wf >= 0 && facet_rule(u32(wf), i, j, d, diw, djw);
}

__device__ __forceinline__ const float* shfl_ptr(const float* p, int src) {
    const u64 v = reinterpret_cast<u64>(p);
    const u32 lo = __shfl(u32(v), src), hi = __shfl(u32(v >> 32), src);
    return reinterpret_cast<const float*>((u64(hi) << 32) | lo);
}

// Apparent flags of up to 64 edges at once, one per lane (whole wave). Each
// lane passes its edge's cloud: mat (its full n x n matrix, row stride ld),
// bmin (its block minima, nb per row), n, and `valid` = the lane holds an edge.
// Phase 1: the lanes walk their top TOPK candidates w = n-1, n-2, ... in
// step; d(i, w) is read in row i (one address for the lanes of one row) and
// d(j, w) as mat[w * n + j], which the full layout holds bit for bit equal to
// d(w, j), so consecutive edges of one row read one coalesced segment.
// Phase 2: each edge still open is searched below its top TOPK by the whole
// wave with the block filter (edge_is_apparent).
template <u32 TOPK>
__device__ __forceinline__ bool edges_apparent_lanes(const float* __restrict__ mat, const float* __restrict__ bmin,
                                                     u32 n, u32 ld, u32 nb, bool valid, u32 i, u32 j,
                                                     float d, u32 lane) {
    bool done = !valid, ap = false;
    for (u32 k = 0; k < TOPK; ++k) {
        if (!done) {
            if (k >= n) {
                done = true;                         // every w examined: no cofacet
            } else {
                const u32 w = n - 1 - k;
                if (w != i && w != j) {
                    const float a = mat[u64(i) * ld + w];
                    const float c = mat[u64(w) * ld + j];
                    if (fmax_ref(a, c) <= d) {
                        ap = facet_rule(w, i, j, d, a, c);
                        done = true;
                    }
                }
            }
        }
        if (!__ballot(!done)) break;
    }
    u64 open = __ballot(!done && n > TOPK);
    while (open) {
        const int src = __builtin_ctzll(open);
        open &= open - 1;
        const u32 ei = __shfl(i, src), ej = __shfl(j, src), en = __shfl(n, src), enb = __shfl(nb, src);
        const u32 eld = __shfl(ld, src);
        const float ed = __shfl(d, src);
        const float* em = shfl_ptr(mat, src);
        const float* eb = shfl_ptr(bmin, src);
        const bool r = edge_is_apparent(em + u64(ei) * eld, em + u64(ej) * eld, eb + u64(ei) * enb,
                                        eb + u64(ej) * enb, en, enb, ei, ej, ed, lane, en - TOPK);
        if (int(lane) == src) ap = r;
    }
    return ap;
}

}  // namespace ttg
