// SPDX-License-Identifier: MIT
// mst.hip — H0 of the Rips filtration of a point cloud: the minimum spanning
// tree of the complete graph by Boruvka rounds.
//
// Replaces mst::h0's Prim loop (include/turbotops/mst.hpp:25-99). The H0
// deaths are the MST edge weights, a multiset every MST shares, so the
// sorted deaths equal Prim's whatever the algorithm.
//
// Edge order: (key, min id, max id), a strict total order, so the MST is
// unique and Boruvka never closes a cycle. key is the distance for L1, L-inf
// and Lp, and for L2 the fma sum before the square root: sqrtf is monotone, so
// that order refines the distance order and the tree is an MST for the
// distances too; the deaths are sqrtf(key) of its edges.
//
// Spatial order (once): points are sorted by a Morton-style code of their
// grid cell and copied in that order; Boruvka runs on the sorted positions.
// Any strict edge order gives a minimum spanning tree, and every MST has the
// same multiset of weights, so the sorted deaths do not depend on the labels.
// Blocks of 64 consecutive points and super-blocks of 64 blocks carry an
// axis-aligned box (the min and max of their members' coordinates).
//
// Per round:
//   todo        vertices that need a fresh search: all in round 0, afterwards
//               those whose previous best has joined their own component (the
//               outside set only shrinks, so any other keeps its edge)
//   comp tags   per block and super-block: the one component of all its
//               members, or "mixed"
//   nearest     one lane per todo vertex finds its smallest (key, j) edge into
//               another component, in two passes: first within the wave's own
//               super-block only, which with the kept vertices' edges gives
//               each component an upper bound of its smallest edge; then over
//               everything, starting from that bound (a vertex with nothing at
//               or below it cannot hold its component's smallest edge; it
//               records none and is searched again next round). Super-blocks, then blocks, are visited from
//               the wave's own block outward in sorted order; one is skipped
//               when it lies wholly in the vertex's component, or when the key
//               to the nearest point of its box exceeds the best so far. The
//               float key is monotone in each coordinate difference (the
//               differences, squares, fused multiply-adds, sums and maxima all
//               round monotonically) and no member is nearer than the box's
//               nearest point in any coordinate, so that key is a lower bound
//               of every member's key: skipping never changes the result. For
//               Lp the bound is the L-inf distance, which the computed Lp
//               distance never goes below (m * pow(s, 1/p) with s >= 1).
//   comp_min    per component, atomicMin of the key bits, then atomicMin of
//               (min id << 32 | max id) among the vertices that reach it
//   hook        each component points at the one across its edge; of a
//               mutual pair the smaller root stays a root; each chosen edge is
//               recorded once
//   jump        pointer jumping until every label names its root
#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "pairs.hpp"
#include "frontend.hpp"

namespace ttg {

namespace {

constexpr u32 MB = 256;
constexpr u32 PB = 64;          // points per block (one wave)
constexpr u32 SBB = 64;         // blocks per super-block
constexpr u32 MIXED = 0xffffffffu;

// Key from q to the nearest point of the box [lo, hi] (see the file comment):
// the same arithmetic as pdist<V, DIM, true> on the clamped coordinates for
// L2 and L1, the largest coordinate difference for L-inf and Lp.
template <int V, int DIM>
__device__ __forceinline__ float box_key(const float* __restrict__ q, const float* __restrict__ lo,
                                         const float* __restrict__ hi, u32 dim_rt) {
    const u32 dim = DIM > 0 ? u32(DIM) : dim_rt;
    float s = 0.0f;
#pragma unroll
    for (u32 k = 0; k < dim; ++k) {
        const float z = __builtin_fminf(__builtin_fmaxf(q[k], lo[k]), hi[k]);
        const float dz = __builtin_fabsf(z - q[k]);
        if constexpr (V == MV_L2) s = __builtin_fmaf(dz, dz, s);
        else if constexpr (V == MV_L1) s += dz;
        else s = __builtin_fmaxf(s, dz);
    }
    return s;
}

// One lane per todo vertex (a sorted position): its smallest (key, j) edge
// to another component, with blocks skipped as described above. Every
// decision to scan a block is a ballot over the wave, and the block's points
// are wave-uniform (scalar loads), so a scanned block costs its 64 pairs once
// per wave.
template <int V, int DIM>
__global__ __launch_bounds__(MB) void mst_nearest(const float* __restrict__ xs, u32 n, u32 dim_rt,
                                                  MetricArgs ma, const u32* __restrict__ comp,
                                                  const u32* __restrict__ todo, const u32* __restrict__ d_ntodo,
                                                  const float* __restrict__ box, const u32* __restrict__ bcomp,
                                                  const float* __restrict__ sbox,
                                                  const u32* __restrict__ sbcomp, u32 nblk,
                                                  float* __restrict__ best_key, u32* __restrict__ best_j,
                                                  unsigned long long* __restrict__ pairs,
                                                  const u32* __restrict__ bound, int local,
                                                  float* __restrict__ lower) {
    constexpr u32 DS = DIM > 0 ? DIM : 1;
    const u32 dim = DIM > 0 ? u32(DIM) : dim_rt;
    const u32 ntodo = *d_ntodo;   // the grid covers every vertex; the todo count is read here
    if (ntodo == 0) return;       // grid-uniform
    const u32 t = blockIdx.x * MB + threadIdx.x;
    if (blockIdx.x * MB >= ntodo) return;   // block-uniform: no todo vertex in this block
    const bool live = t < ntodo;
    const u32 i = live ? todo[t] : todo[ntodo - 1];
    const u32 ci = comp[i];
    // bounded pass: lower[i] never exceeds i's smallest outside key (that key
    // only grows from round to round, as the outside set shrinks), so a vertex
    // whose lower[i] is above its component's bound cannot hold the
    // component's smallest edge and sits this search out
    const bool search = live && !(bound && lower[i] > __uint_as_float(bound[ci]));
    float qr[DS];
    const float* __restrict__ q = xs + u64(i) * dim;
    if constexpr (DIM > 0) {
#pragma unroll
        for (int k = 0; k < DIM; ++k) qr[k] = q[k];
        q = qr;
    }
    // bounded search: only edges with key <= the component's bound can be
    // its smallest, so the search starts there (ties admitted: bj = ~0)
    float bk = bound ? __uint_as_float(bound[ci]) : INFINITY;
    u32 bj = 0xffffffffu;
    const u32 nsb = (nblk + SBB - 1) / SBB;
    const u32 home = __builtin_amdgcn_readfirstlane(i) / PB;   // the block of the wave's first vertex
    const u32 hsb = home / SBB;
    const u32 nlive = u32(__popcll(__ballot(search)));
    u64 measured = 0;   // pairs this wave evaluates
    const u32 nsu = local ? 1u : nsb;   // local: the wave's own super-block only
    for (u32 su = 0; su < nsu; ++su) {
        u32 sb = hsb + su;
        if (sb >= nsb) sb -= nsb;
        {
            const float* bx = sbox + u64(sb) * 2 * dim;
            const float lb = box_key<V, DIM>(q, bx, bx + dim, dim);
            if (!__ballot(search && sbcomp[sb] != ci && lb <= bk)) continue;
        }
        const u32 b0 = sb * SBB, len = min(SBB, nblk - b0);
        const u32 start = sb == hsb ? home - b0 : 0u;
        for (u32 bu = 0; bu < len; ++bu) {
            u32 b = start + bu;
            if (b >= len) b -= len;
            b += b0;
            {
                const float* bx = box + u64(b) * 2 * dim;
                const float lb = box_key<V, DIM>(q, bx, bx + dim, dim);
                if (!__ballot(search && bcomp[b] != ci && lb <= bk)) continue;
            }
            const u32 j1 = u32(min<u64>(n, u64(b + 1) * PB));
            measured += u64(nlive) * (j1 - b * PB);
            for (u32 j = b * PB; j < j1; ++j) {
                const u32 cj = comp[j];
                float kv;
                if constexpr (DIM > 0) {
                    float c[DIM];
#pragma unroll
                    for (int k = 0; k < DIM; ++k) c[k] = xs[u64(j) * DIM + k];
                    kv = pdist<V, DIM, true>(c, q, dim, ma);
                } else {
                    kv = pdist<V, 0, true>(xs + u64(j) * dim, q, dim, ma);
                }
                // (key, j) order: for a fixed i this is the edge order (key, min, max)
                if (cj != ci && (kv < bk || (kv == bk && j < bj))) {
                    bk = kv;
                    bj = j;
                }
            }
        }
    }
    if (live) {
        // nothing at or below the bound (or sat out): this vertex cannot hold
        // its component's smallest edge, and it is searched again next round
        const bool found = search && bj != 0xffffffffu;
        best_key[i] = found ? bk : INFINITY;
        best_j[i] = found ? bj : 0xffffffffu;
        // its next smallest outside key is at least the one found, and above
        // the bound it searched against
        if (bound && search) lower[i] = found ? bk : __uint_as_float(bound[ci]);
    }
    if ((threadIdx.x & 63) == 0 && measured) atomicAdd(pairs, (unsigned long long)measured);
}

template <int V, int DIM>
struct NearestLaunch {
    static void run(u32 blocks, hipStream_t s, const float* xs, u32 n, u32 dim, MetricArgs ma, const u32* comp,
                    const u32* todo, const u32* ntodo, const float* box, const u32* bcomp, const float* sbox,
                    const u32* sbcomp, u32 nblk, float* bk, u32* bj, unsigned long long* pairs,
                    const u32* bound, int local, float* lower) {
        mst_nearest<V, DIM><<<blocks, MB, 0, s>>>(xs, n, dim, ma, comp, todo, ntodo, box, bcomp, sbox, sbcomp,
                                                  nblk, bk, bj, pairs, bound, local, lower);
    }
};

// One wave per block (or super-block): the one component of its members
// (points' comp, or blocks' tags), or MIXED.
__global__ __launch_bounds__(MB) void mst_tags(const u32* __restrict__ src, u32 count, u32* __restrict__ out) {
    const u32 lane = threadIdx.x & 63;
    const u32 nbox = (count + 63) / 64;
    for (u32 b = (blockIdx.x * MB + threadIdx.x) / 64; b < nbox; b += gridDim.x * (MB / 64)) {
        const u32 m = min(64u, count - b * 64);
        const u32 c = src[b * 64 + (lane < m ? lane : 0u)];
        const u32 c0 = __builtin_amdgcn_readfirstlane(c);
        const bool one = c0 != MIXED && __ballot(c != c0) == 0;   // every lane votes
        if (lane == 0) out[b] = one ? c0 : MIXED;
    }
}

__device__ __forceinline__ u32 lane_rank_u64(u64 mask) {
    return __builtin_amdgcn_mbcnt_hi(u32(mask >> 32), __builtin_amdgcn_mbcnt_lo(u32(mask), 0u));
}

// Vertices that need a fresh search: every vertex in round 0, afterwards
// those whose best neighbour has joined their own component. Each wave
// appends its vertices in lane order (one atomic per wave), so a wave of the
// search takes neighbouring positions.
__global__ __launch_bounds__(MB) void mst_todo(const u32* __restrict__ comp, const u32* __restrict__ best_j, u32 n,
                                               int first, u32* __restrict__ todo, u32* __restrict__ ntodo) {
    __shared__ u32 wcount[MB / 64], sbase;
    const u32 lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    // one atomic per workgroup and step (one per wave serialised on the counter)
    for (u32 b0 = blockIdx.x * MB; b0 < n; b0 += gridDim.x * MB) {   // block-uniform
        const u32 v = b0 + threadIdx.x;
        const bool need = v < n && (first || best_j[v] == 0xffffffffu || comp[best_j[v]] == comp[v]);
        const u64 mask = __ballot(need);
        if (lane == 0) wcount[w] = u32(__popcll(mask));
        __syncthreads();
        if (threadIdx.x == 0) {
            u32 tot = 0;
            for (u32 q = 0; q < MB / 64; ++q) tot += wcount[q];
            sbase = tot ? atomicAdd(ntodo, tot) : 0u;
        }
        __syncthreads();
        u32 off = sbase;
        for (u32 q = 0; q < w; ++q) off += wcount[q];
        if (need) todo[off + lane_rank_u64(mask)] = v;
        __syncthreads();
    }
}

// Per component, the smallest key of its vertices. Neighbouring positions
// mostly share a component once components are large, and atomics on one
// address serialise (one per wave cost 1.4 ms per launch at n = 1e6 with a few
// components left), so each workgroup takes a range of 4096 positions: a wave
// keeps a running minimum over its 16 consecutive 64-vertex chunks while they
// all hold one component, the four waves' runs are merged, and one atomic per
// component and range remains. Mixed chunks fall back to one atomic per vertex.
constexpr u32 CK_RANGE = 4096;

__global__ __launch_bounds__(MB) void mst_comp_key(const u32* __restrict__ comp, const float* __restrict__ best_key,
                                                   u32 n, u32* __restrict__ cw) {
    __shared__ u32 sc[MB / 64], sk[MB / 64];
    const u32 lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    constexpr u32 CHUNKS = CK_RANGE / MB;   // chunks of 64 per wave
    for (u64 r0 = u64(blockIdx.x) * CK_RANGE; r0 < n; r0 += u64(gridDim.x) * CK_RANGE) {
        u32 run_c = MIXED, run_k = 0xffffffffu;   // wave-uniform
        for (u32 t = 0; t < CHUNKS; ++t) {
            const u64 v = r0 + u64(w * CHUNKS + t) * 64 + lane;
            const bool in = v < n;
            if (!__ballot(in)) break;   // past the end: uniform
            const u32 c = in ? comp[v] : 0u;
            const u32 key = in ? __float_as_uint(best_key[v]) : 0xffffffffu;
            const u32 c0 = __builtin_amdgcn_readfirstlane(c);   // lane 0 is in range
            if (__ballot(in && c != c0) == 0) {
                u32 km = key;
                for (int off = 32; off > 0; off >>= 1) km = min(km, u32(__shfl_xor(int(km), off)));
                if (c0 != run_c) {
                    if (run_c != MIXED && lane == 0) atomicMin(cw + run_c, run_k);
                    run_c = c0;
                    run_k = km;
                } else {
                    run_k = min(run_k, km);
                }
            } else {
                // a chunk across components: one minimum and one atomic per
                // component present (per-lane atomics here serialised on the
                // few large components of the late rounds)
                u64 rem = __ballot(in);
                while (rem) {
                    const u32 cq = __builtin_amdgcn_readlane(c, int(__builtin_ctzll(rem)));
                    const bool mine = in && c == cq;
                    u32 km = mine ? key : 0xffffffffu;
                    for (int off = 32; off > 0; off >>= 1) km = min(km, u32(__shfl_xor(int(km), off)));
                    if (cq == run_c) run_k = min(run_k, km);
                    else if (lane == 0) atomicMin(cw + cq, km);
                    rem &= ~__ballot(mine);
                }
            }
        }
        if (lane == 0) {
            sc[w] = run_c;
            sk[w] = run_k;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (u32 q = 0; q < MB / 64; ++q) {
                if (sc[q] == MIXED) continue;
                u32 kk = sk[q];
                for (u32 q2 = q + 1; q2 < MB / 64; ++q2)
                    if (sc[q2] == sc[q]) {
                        kk = min(kk, sk[q2]);
                        sc[q2] = MIXED;
                    }
                atomicMin(cw + sc[q], kk);
            }
        }
        __syncthreads();
    }
}

u32 comp_key_blocks(u32 n) { return u32(std::max<u64>(1, std::min<u64>((u64(n) + CK_RANGE - 1) / CK_RANGE, 16384))); }

__global__ void mst_comp_pair(const u32* __restrict__ comp, const float* __restrict__ best_key,
                              const u32* __restrict__ best_j, u32 n, const u32* __restrict__ cw,
                              unsigned long long* __restrict__ cp) {
    for (u32 v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
        const u32 c = comp[v];
        if (__float_as_uint(best_key[v]) != cw[c] || best_j[v] == 0xffffffffu) continue;
        const u32 j = best_j[v];
        const u64 a = min(v, j), b = max(v, j);
        atomicMin(cp + c, (unsigned long long)((a << 32) | b));
    }
}

__global__ void mst_hook(const u32* __restrict__ comp, u32 n, const u32* __restrict__ cw,
                         const unsigned long long* __restrict__ cp, u32* __restrict__ parent,
                         float* __restrict__ out_keys, u32* __restrict__ out_count) {
    for (u32 c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
        if (comp[c] != c) continue;            // not a root
        const u64 e = cp[c];
        if (e == ~0ull) continue;              // no outgoing edge (single component)
        const u32 u = u32(e >> 32), v = u32(e);
        const u32 cu = comp[u], cv = comp[v];
        const u32 other = cu == c ? cv : cu;
        const bool mutual = cp[other] == e && cw[other] == cw[c];
        if (mutual && c < other) {
            out_keys[atomicAdd(out_count, 1u)] = __uint_as_float(cw[c]);
        } else {
            parent[c] = other;
            if (!mutual) out_keys[atomicAdd(out_count, 1u)] = __uint_as_float(cw[c]);
        }
    }
}

__global__ void mst_jump(u32* __restrict__ parent, u32 n, u32* __restrict__ changed) {
    for (u32 c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
        const u32 p = parent[c];
        const u32 gp = parent[p];
        if (gp != p) {
            parent[c] = gp;
            *changed = 1u;
        }
    }
}

__global__ void mst_relabel(u32* __restrict__ comp, const u32* __restrict__ parent, u32 n,
                            u32* __restrict__ cw, unsigned long long* __restrict__ cp) {
    for (u32 v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
        comp[v] = parent[comp[v]];
        cw[v] = 0xffffffffu;
        cp[v] = ~0ull;
    }
}

template <int V>
__global__ void mst_finish(float* __restrict__ keys, u32 m, u64* __restrict__ sortkeys) {
    for (u32 t = blockIdx.x * blockDim.x + threadIdx.x; t < m; t += gridDim.x * blockDim.x)
        sortkeys[t] = __float_as_uint(key_to_dist<V>(keys[t]));
}

__global__ void mst_unpack(const u64* __restrict__ sortkeys, u32 m, float* __restrict__ out) {
    for (u32 t = blockIdx.x * blockDim.x + threadIdx.x; t < m; t += gridDim.x * blockDim.x)
        out[t] = __uint_as_float(u32(sortkeys[t]));
}

__global__ void iota_u32(u32* __restrict__ p, u32 n) {
    for (u32 v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) p[v] = v;
}

u32 blocks_of(u64 count) {
    return u32(std::max<u64>(1, std::min<u64>((count + MB - 1) / MB, 16384)));
}

}  // namespace


struct MstStats {
    u32 rounds = 0;
    u64 scanned = 0;    // vertex searches
    u64 pairs = 0;      // point pairs measured by them
};

namespace {

// Sub-buffers of one device region, handed out in order at 256-byte
// alignment. With base == nullptr it only measures (the bytes a run needs).
struct Carve {
    char* base = nullptr;
    u64 cap = ~0ull, used = 0;
    template <typename T>
    T* get(u64 count) {
        const u64 off = (used + 255) & ~u64(255);
        const u64 need = count * sizeof(T);
        if (off + need > cap) fail(TT_ELENGTH, "mst: scratch region too small");
        used = off + need;
        return base ? reinterpret_cast<T*>(base + off) : nullptr;
    }
};

// Every device buffer a Boruvka run on n vertices in R^dim uses, carved from
// `c` in a fixed order (a measuring Carve gives the total).
struct MstBuffers {
    u32 nblk, nsb;
    u32 *comp, *parent, *cw, *bj, *todo, *scal, *bcomp, *sbcomp, *ext;
    unsigned long long* cp;
    float *bk, *keys, *xs, *box, *sbox, *lower;
    u64 *sk, *alt, *sort_scratch;
    MstBuffers(Carve& c, u32 n, u32 dim) {
        nblk = (n + PB - 1) / PB;
        nsb = (nblk + SBB - 1) / SBB;
        comp = c.get<u32>(n);
        parent = c.get<u32>(n);
        cw = c.get<u32>(n);
        cp = c.get<unsigned long long>(n);
        bk = c.get<float>(n);
        lower = c.get<float>(n);
        bj = c.get<u32>(n);
        todo = c.get<u32>(n);
        scal = c.get<u32>(8);      // [0] ntodo, [1] changed, [2] out count, [3] n - 1, [4] n, [6..7] pairs
        keys = c.get<float>(n);
        sk = c.get<u64>(n);
        alt = c.get<u64>(n);
        sort_scratch = c.get<u64>(radix_scratch_entries(n));
        xs = c.get<float>(u64(n) * dim);
        box = c.get<float>(u64(nblk) * 2 * dim);
        sbox = c.get<float>(u64(nsb) * 2 * dim);
        bcomp = c.get<u32>(nblk);
        sbcomp = c.get<u32>(nsb);
        ext = c.get<u32>(spatial_ext_entries(dim));
    }
};

}  // namespace

u64 mst_scratch_bytes(u32 n, u32 dim) {
    Carve c;
    MstBuffers b(c, std::max<u32>(n, 2), std::max<u32>(dim, 1));
    (void)b;
    return c.used;
}

// Sorted MST edge weights (n - 1 floats, device) of the complete graph, with
// every temporary carved from the device region [base, base + cap).
void mst_deaths_in(const float* d_x, u32 n, u32 dim, int kind, float p, float* d_deaths, char* base, u64 cap,
                   hipStream_t s, MstStats* st) {
    if (n < 2) return;
    const MetricVariant v = variant_of(kind, p);
    const MetricArgs ma = metric_args(kind, p);
    Carve c{base, cap, 0};
    const MstBuffers B(c, n, dim);
    u32 *comp = B.comp, *parent = B.parent, *cw = B.cw, *bj = B.bj, *todo = B.todo, *scal = B.scal;
    unsigned long long* cp = B.cp;
    float *bk = B.bk, *keys = B.keys;
    iota_u32<<<blocks_of(n), MB, 0, s>>>(comp, n);
    iota_u32<<<blocks_of(n), MB, 0, s>>>(parent, n);
    fill_device_u32(cw, n, 0xffffffffu, s);
    fill_device_u32(reinterpret_cast<u32*>(cp), 2 * u64(n), 0xffffffffu, s);
    fill_device_u32(bj, n, 0xffffffffu, s);
    fill_device_u32(reinterpret_cast<u32*>(B.lower), n, 0u, s);   // +0.0f: no bound known yet
    fill_device_u32(scal, 3, 0u, s);
    fill_device_u32(scal + 3, 1, n - 1, s);
    fill_device_u32(scal + 4, 1, n, s);
    fill_device_u32(scal + 6, 2, 0u, s);   // [6..7]: pairs evaluated (u64)
    TTG_LAUNCH_CHECK();

    // spatial order; the sorted copy B.xs is what the rounds read
    spatial_order(d_x, n, dim, scal + 4, B.sk, B.alt, B.sort_scratch, B.ext, B.xs, nullptr, s);
    spatial_boxes(B.xs, n, PB, dim, false, B.box, s);
    spatial_boxes(B.box, B.nblk, SBB, dim, true, B.sbox, s);
    u32 edges = 0;
    for (u32 round = 0; edges < n - 1; ++round) {
        if (round > 64) fail(TT_EDEVICE, "mst: Boruvka did not converge");
        fill_device_u32(scal, 1, 0u, s);
        mst_todo<<<blocks_of(n), MB, 0, s>>>(comp, bj, n, round == 0, todo, scal);
        // the searches read the todo count on the device (their grid covers n)
        mst_tags<<<blocks_of(u64(B.nblk) * 64), MB, 0, s>>>(comp, n, B.bcomp);
        mst_tags<<<blocks_of(u64(B.nsb) * 64), MB, 0, s>>>(B.bcomp, B.nblk, B.sbcomp);
        unsigned long long* pairs = reinterpret_cast<unsigned long long*>(scal + 6);
        const u32 nblocks = u32((u64(n) + MB - 1) / MB);
        // 1: each todo vertex searches its own super-block only; with the kept
        //    vertices' edges that bounds every component's smallest edge from
        //    above (cw)
        dispatch_metric_dim<NearestLaunch>(v, dim, nblocks, s, B.xs, n, dim, ma, comp, todo,
                                           static_cast<const u32*>(scal), B.box, B.bcomp, B.sbox, B.sbcomp, B.nblk,
                                           bk, bj, pairs, static_cast<const u32*>(nullptr), 1, B.lower);
        mst_comp_key<<<comp_key_blocks(n), MB, 0, s>>>(comp, bk, n, cw);
        // 2: the full search, from that bound
        dispatch_metric_dim<NearestLaunch>(v, dim, nblocks, s, B.xs, n, dim, ma, comp, todo,
                                           static_cast<const u32*>(scal), B.box, B.bcomp, B.sbox, B.sbcomp, B.nblk,
                                           bk, bj, pairs, static_cast<const u32*>(cw), 0, B.lower);
        fill_device_u32(cw, n, 0xffffffffu, s);
        TTG_LAUNCH_CHECK();
        mst_comp_key<<<comp_key_blocks(n), MB, 0, s>>>(comp, bk, n, cw);
        mst_comp_pair<<<blocks_of(n), MB, 0, s>>>(comp, bk, bj, n, cw, cp);
        mst_hook<<<blocks_of(n), MB, 0, s>>>(comp, n, cw, cp, parent, keys, scal + 2);
        TTG_LAUNCH_CHECK();
        // pointer jumping, three steps per check of the flag
        for (int it = 0; it < 32; ++it) {
            fill_device_u32(scal + 1, 1, 0u, s);
            for (int r = 0; r < 3; ++r) mst_jump<<<blocks_of(n), MB, 0, s>>>(parent, n, scal + 1);
            u32 changed = 0;
            TTG_HIP(hipMemcpyAsync(&changed, scal + 1, sizeof(u32), hipMemcpyDeviceToHost, s));
            TTG_HIP(hipStreamSynchronize(s));
            if (!changed) break;
        }
        mst_relabel<<<blocks_of(n), MB, 0, s>>>(comp, parent, n, cw, cp);
        TTG_LAUNCH_CHECK();
        u32 counts[3] = {0, 0, 0};   // todo count, (jump flag), edges so far
        TTG_HIP(hipMemcpyAsync(counts, scal, sizeof(counts), hipMemcpyDeviceToHost, s));
        TTG_HIP(hipStreamSynchronize(s));
        edges = counts[2];
        if (st) {
            st->rounds = round + 1;
            st->scanned += counts[0];
        }
    }
    if (edges != n - 1) fail(TT_EDEVICE, "mst: spanning tree has the wrong edge count");
    if (st) {
        unsigned long long pairs = 0;
        TTG_HIP(hipMemcpyAsync(&pairs, scal + 6, sizeof(pairs), hipMemcpyDeviceToHost, s));
        TTG_HIP(hipStreamSynchronize(s));
        st->pairs = pairs;
    }
    switch (v) {
        case MV_L2: mst_finish<MV_L2><<<blocks_of(n), MB, 0, s>>>(keys, n - 1, B.sk); break;
        default: mst_finish<MV_L1><<<blocks_of(n), MB, 0, s>>>(keys, n - 1, B.sk); break;
    }
    TTG_LAUNCH_CHECK();
    const u64* sorted = radix_sort_u64_counted(B.sk, n - 1, scal + 3, 0, 31, B.alt, B.sort_scratch, s);
    mst_unpack<<<blocks_of(n), MB, 0, s>>>(sorted, n - 1, d_deaths);
    TTG_LAUNCH_CHECK();
}

// The same with the region from a Workspace.
void mst_deaths_dev(const float* d_x, u32 n, u32 dim, int kind, float p, float* d_deaths, Workspace& ws,
                    hipStream_t s, MstStats* st) {
    if (n < 2) return;
    const u64 bytes = mst_scratch_bytes(n, dim);
    mst_deaths_in(d_x, n, dim, kind, p, d_deaths, ws.get<char>(bytes), bytes, s, st);
}

// ---- the no-allocation form ---------------------------------------------
// One scratch region per device for tt_mst_deaths_dev_noalloc: a file-scope
// device array, so a call allocates nothing. Host code never passes the
// array's name to a kernel (on the host it is only a stand-in for the
// symbol); it passes the device address hipGetSymbolAddress returns, looked
// up once per device with that device current.
constexpr u32 kMstNoAllocMaxN = 1u << 20;
constexpr u64 kMstScratchBytes = u64(128) << 20;   // n (56.6 + 4.125 dim) bytes: dim <= 17 at n = 2^20

static __device__ __align__(256) unsigned char scratch[kMstScratchBytes];

namespace {

std::mutex g_scratch_mu;
std::vector<char*> g_scratch_ptr;   // per device ordinal; nullptr until looked up

char* scratch_ptr() {
    int dev = 0;
    TTG_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr.size() <= size_t(dev)) g_scratch_ptr.resize(size_t(dev) + 1, nullptr);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            fail(TT_EDEVICE, std::string("hip: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<char*>(p);
    }
    return g_scratch_ptr[dev];
}

}  // namespace

}  // namespace ttg

using namespace ttg;

extern "C" {

tt_status tt_mst_deaths(const float* pts, uint32_t n, uint32_t dim, int kind, float p,
                        float* deaths) {
    return guarded([&] {
        check_metric(kind, p);
        check_points(pts, u64(n) * dim, "tt_mst_deaths");
        for (u64 t = 0; t < u64(n) * dim; ++t)
            if (!std::isfinite(pts[t])) fail(TT_EINVAL, "mst: non-finite coordinate");
        if (n < 2) return;
        if (dim == 0) fail(TT_EINVAL, "tt_mst_deaths: dim must be >= 1");
        if (!deaths) fail(TT_EINVAL, "tt_mst_deaths: null deaths");
        hipStream_t s = nullptr;
        Workspace ws;
        float* dx = ws.get<float>(u64(n) * dim);
        float* dd = ws.get<float>(n - 1);
        TTG_HIP(hipMemcpyAsync(dx, pts, u64(n) * dim * sizeof(float), hipMemcpyHostToDevice, s));
        mst_deaths_dev(dx, n, dim, kind, p, dd, ws, s, nullptr);
        TTG_HIP(hipMemcpyAsync(deaths, dd, u64(n - 1) * sizeof(float), hipMemcpyDeviceToHost, s));
        TTG_HIP(hipStreamSynchronize(s));
    });
}

tt_status tt_mst_deaths_dev(const float* d_pts, uint32_t n, uint32_t dim, int kind, float p,
                            float* d_deaths, uint32_t* rounds, uint64_t* scans, uint64_t* pairs,
                            void* stream) {
    return guarded([&] {
        check_metric(kind, p);
        if (n >= 2 && (!d_pts || !d_deaths)) fail(TT_EINVAL, "tt_mst_deaths_dev: null buffer");
        if (n >= 2 && dim == 0) fail(TT_EINVAL, "tt_mst_deaths_dev: dim must be >= 1");
        hipStream_t s = as_stream(stream);
        Workspace ws;
        MstStats st;
        mst_deaths_dev(d_pts, n, dim, kind, p, d_deaths, ws, s, &st);
        TTG_HIP(hipStreamSynchronize(s));
        if (rounds) *rounds = st.rounds;
        if (scans) *scans = st.scanned;
        if (pairs) *pairs = st.pairs;
    });
}

tt_status tt_mst_deaths_dev_noalloc(const float* d_pts, uint32_t n, uint32_t dim, int kind, float p,
                                    float* d_deaths, void* stream) {
    return guarded([&] {
        check_metric(kind, p);
        if (n > kMstNoAllocMaxN)
            fail(TT_ELENGTH, "tt_mst_deaths_dev_noalloc: n exceeds " + std::to_string(kMstNoAllocMaxN) +
                                 " points (the static scratch region)");
        if (n >= 2 && (!d_pts || !d_deaths)) fail(TT_EINVAL, "tt_mst_deaths_dev_noalloc: null buffer");
        if (n >= 2 && dim == 0) fail(TT_EINVAL, "tt_mst_deaths_dev_noalloc: dim must be >= 1");
        if (n < 2) return;
        const u64 bytes = mst_scratch_bytes(n, dim);
        if (bytes > kMstScratchBytes) fail(TT_ELENGTH, "tt_mst_deaths_dev_noalloc: scratch region too small");
        hipStream_t s = as_stream(stream);
        mst_deaths_in(d_pts, n, dim, kind, p, d_deaths, scratch_ptr(), kMstScratchBytes, s, nullptr);
        TTG_HIP(hipStreamSynchronize(s));
    });
}

}  // extern "C"
