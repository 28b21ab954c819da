// SPDX-License-Identifier: MIT
// apparent.hpp — the zero-apparent test of edges on a full distance matrix,
// 64 edges per wave: a lane-parallel walk over the top candidates, then a
// block-minimum filtered search by 16-lane groups for the edges still open.
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
// bmin[j][b] <= d, so one pass over the block minima picks the blocks worth
// reading, highest first, and only those are read. The filter is
// conservative (a block's minimum includes w = i and w = j), so the first w
// found is exactly the plain walk's.
#pragma once

#include "common.hpp"

namespace ttg {

constexpr u32 BLK = 64;
constexpr u32 TOPK = 64;   // candidates the lanes walk together before the grouped search

__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
    for (u32 off = 32; off > 0; off >>= 1) v = fminf(v, __shfl_xor(v, off));
    return v;
}

// bmin_row[b] for every block b of one row (whole wave; lane = lane id).
// Eight blocks per step: their loads are issued together (one 256-byte
// segment at a time left the pass latency-bound), and lane u stores the
// minimum of the step's u-th block.
__device__ __forceinline__ void block_min_row(const float* __restrict__ row, u32 n, u32 nb,
                                              float* __restrict__ bmin_row, u32 lane) {
    constexpr u32 STEP = 8;
    for (u32 b0 = 0; b0 < nb; b0 += STEP) {
        float v[STEP];
#pragma unroll
        for (u32 u = 0; u < STEP; ++u) {
            const u32 w = (b0 + u) * BLK + lane;
            v[u] = (b0 + u < nb && w < n) ? row[w] : INFINITY;
        }
        float mine = INFINITY;
#pragma unroll
        for (u32 u = 0; u < STEP; ++u) {
            const float m = wave_min(v[u]);
            if (lane == u) mine = m;
        }
        if (lane < STEP && b0 + lane < nb) bmin_row[b0 + lane] = mine;
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

// The boundary engine's ready test (host/zerodiff.hpp) for the triangle
// {lo, hi, w} with max(d(lo, w), d(hi, w)) <= D: the edge (lo, hi) of
// diameter D is its youngest face in (diameter, comb index) order.
__device__ __forceinline__ bool witness_rule(u32 w, u32 lo, u32 hi, float D, float dlo, float dhi) {
    if (w < lo) return true;                    // both other faces come first
    if (w < hi) return dhi < D;                 // face {w, hi} comes later
    return dlo < D && dhi < D;                  // faces {lo, w} and {hi, w} come later
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
// Phase 2: each edge still open is searched below its top TOPK with the block
// filter: the block minima of rows i and j pick the blocks that can hold a
// w with max(d(i,w), d(j,w)) <= d, highest first, and a ballot over each such
// block finds the first cofacet. The filter is conservative (a block's minimum
// includes w = i and w = j), so the first w found is exactly the plain walk's.
//
// WITNESS = true: the same search for the boundary engine's ready witness of
// the edge (j, i), j < i: any w with max(d(i, w), d(j, w)) <= d that passes
// witness_rule; the result is whether one exists.
template <u32 TOPK, bool WITNESS = false>
__device__ __forceinline__ bool edges_apparent_lanes(const float* __restrict__ mat, const float* __restrict__ bmin,
                                                     u32 n, u32 ld, u32 nb, bool valid, u32 i, u32 j,
                                                     float d, u32 lane) {
    bool done = !valid, ap = false;
    // Phase 1 in steps of P1 candidates: a lane loads the step's d(i, w) and
    // d(j, w) together, then examines them in descending w, as the one-by-one
    // walk does (one dependent load pair per candidate left the walk bound by
    // cache latency).
    constexpr u32 P1 = 8;
    static_assert(TOPK % P1 == 0, "phase 1 runs in whole steps");
    for (u32 k0 = 0; k0 < TOPK; k0 += P1) {
        float av[P1], cv[P1];
#pragma unroll
        for (u32 u = 0; u < P1; ++u) {
            const u32 k = k0 + u;
            const u32 w = n - 1 - k;
            const bool in = !done && k < n && w != i && w != j;
            av[u] = in ? mat[u64(i) * ld + w] : INFINITY;
            cv[u] = in ? mat[u64(w) * ld + j] : INFINITY;
        }
#pragma unroll
        for (u32 u = 0; u < P1; ++u) {
            const u32 k = k0 + u;
            if (!done) {
                if (k >= n) {
                    done = true;                         // every w examined: no cofacet
                } else {
                    const u32 w = n - 1 - k;
                    const float a = av[u], c = cv[u];
                    if (w != i && w != j && fmax_ref(a, c) <= d) {
                        if constexpr (WITNESS) {
                            if (witness_rule(w, j, i, d, c, a)) {
                                ap = true;
                                done = true;
                            }
                        } else {
                            ap = facet_rule(w, i, j, d, a, c);
                            done = true;
                        }
                    }
                }
            }
        }
        if (!__ballot(!done)) break;
    }
    // Phase 2, four open edges at a time: lanes 16g .. 16g + 15 search the
    // g-th one, each lane checking 4 of 64 blocks (or of 64 candidates) per
    // step, so four searches overlap their dependent loads.
    u64 open = __ballot(!done && n > TOPK);
    const u32 g = lane >> 4, gl = lane & 15;
    while (open) {
        int src[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            src[t] = open ? __builtin_ctzll(open) : -1;
            open &= open ? open - 1 : 0;
        }
        const int mine = g == 0 ? src[0] : g == 1 ? src[1] : g == 2 ? src[2] : src[3];
        const bool active = mine >= 0;
        const int sl = active ? mine : int(lane);
        const u32 ei = __shfl(i, sl), ej = __shfl(j, sl), en = __shfl(n, sl), enb = __shfl(nb, sl);
        const u32 eld = __shfl(ld, sl);
        const float ed = __shfl(d, sl);
        const float* em = shfl_ptr(mat, sl);
        const float* eb = shfl_ptr(bmin, sl);
        const float* ri = em + u64(ei) * eld;
        const float* rj = em + u64(ej) * eld;
        const float* bi = eb + u64(ei) * enb;
        const float* bj = eb + u64(ej) * enb;
        const u32 w_hi = active ? en - TOPK : 0;           // phase 1 covered w >= w_hi
        int top = active ? int((w_hi - 1) / BLK) : -1;
        int wf = -1;
        float diw = 0.0f, djw = 0.0f;
        while (__ballot(active && wf < 0 && top >= 0)) {
            const bool live = active && wf < 0 && top >= 0;
            u64 cmask = 0;   // bit 16k + t: block top - t - 16k may hold a cofacet
#pragma unroll
            for (u32 k = 0; k < 4; ++k) {
                const int b = top - int(gl) - 16 * int(k);
                bool cand = false;
                if (live && b >= 0) cand = fmax_ref(bi[b], bj[b]) <= ed;
                cmask |= ((__ballot(cand) >> (16 * g)) & 0xffffull) << (16 * k);
            }
            while (__ballot(live && cmask != 0 && wf < 0)) {
                const bool work = live && cmask != 0 && wf < 0;
                const int blk = work ? top - __builtin_ctzll(cmask) : 0;   // highest block first
                if (work) cmask &= cmask - 1;
                u64 hit = 0;     // bit 16k + t: w = 64 blk + 63 - t - 16k passes
                float av[4], cv[4];
#pragma unroll
                for (u32 k = 0; k < 4; ++k) {
                    const int w = blk * int(BLK) + int(BLK - 1) - int(gl) - 16 * int(k);
                    const bool in = work && u32(w) < w_hi;
                    float a = INFINITY, c = INFINITY;
                    if (in) {
                        a = ri[w];
                        c = rj[w];
                    }
                    av[k] = a;
                    cv[k] = c;
                    bool ok = in && u32(w) != ei && u32(w) != ej && fmax_ref(a, c) <= ed;
                    if constexpr (WITNESS) ok = ok && witness_rule(u32(w), ej, ei, ed, c, a);
                    hit |= ((__ballot(ok) >> (16 * g)) & 0xffffull) << (16 * k);
                }
                const int q = hit ? __builtin_ctzll(hit) : 0;     // lowest bit = largest w
                const int kq = q >> 4, from = int(16 * g) + (q & 15);
                float a = 0.0f, c = 0.0f;
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const float ak = __shfl(av[k], from), ck = __shfl(cv[k], from);
                    if (k == kq) {
                        a = ak;
                        c = ck;
                    }
                }
                if (work && hit) {
                    wf = blk * int(BLK) + int(BLK - 1) - (q & 15) - 16 * kq;
                    diw = a;
                    djw = c;
                }
            }
            if (live && wf < 0) top -= 64;
        }
        bool r;
        if constexpr (WITNESS) r = wf >= 0;
        else r = wf >= 0 && facet_rule(u32(wf), ei, ej, ed, diw, djw);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const bool rt = __shfl(int(r), 16 * t) != 0;
            if (int(lane) == src[t]) ap = rt;
        }
    }
    return ap;
}

}  // namespace ttg
