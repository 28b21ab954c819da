// SPDX-License-Identifier: MIT
// columns.hip — the coboundary reducer's per-simplex work above dimension 0
// on the GPU: the columns of each next dimension and, per column, the
// pivot-independent half of the emergent-pair shortcut.
//
// Replaces, on a dense (full layout) or sparse (CSR graph) complex:
//   Reducer::assemble (include/turbotops/rips.hpp:981-1023): every cofacet
//     with a new top vertex of every simplex of the current dimension, up to
//     the threshold; the ones that are pivots of the reduction just done are
//     cleared, the ones in a zero-apparent pair either way
//     (zero_apparent_cofacet_v / zero_apparent_facet_v, :638-665) skipped;
//     the rest are the columns of the next dimension in descending
//     filtration order (d descending, index ascending; sort_desc :776-778);
//   compute_pairs' emergent shortcut (:903-928) without its pivot lookup:
//     the first cofacet of equal diameter in descending index order (the
//     zero-pivot cofacet, :585-605) and whether that cofacet has a
//     zero-apparent facet. The host pairs a column with it at once when it is
//     not yet a pivot and has no such facet.
//
// The tests are stated on vertex sets, not on indices:
//   zpc(U, d)  the largest w outside U with every d(u, w) <= d (on a graph,
//              every (u, w) an edge): the zero-pivot cofacet is U + {w};
//   zpf(U, d)  the first facet in ascending index order — vertex position
//              j = 0, 1, ... removed (descending vertex order) — that keeps an
//              edge of length d; with M the edges of length d, that is the
//              first position outside the intersection of their end points;
//   (U, U + {w}) apparent  iff  zpf(U + {w}) removes w;
//   (F, U) apparent, F = zpf(U) removing r  iff  no w > r outside F has
//              every d(f, w) <= d (r itself qualifies, so zpc(F) >= r).
//
// Pipeline per assemble (one stream):
//   cob_count / scan / cob_emit   every cofacet (index, diameter): one lane
//                                 per simplex, candidates x > top vertex
//   col_flags                     one lane per cofacet: cleared (binary
//                                 search in the sorted pivots), apparent,
//                                 emergent state and cofacet
//   scan / col_scatter            the columns, compacted
//   radix sorts                   index ascending (64-bit words, low then
//                                 high), then diameter descending, stable,
//                                 carrying a permutation
//   col_gather                    the columns in that order
// The cofacets stay in HBM as the next dimension's simplices.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <vector>

#include "frontend.hpp"

namespace ttg {

using u128 = unsigned __int128;

namespace {

constexpr u32 CB = 256;
constexpr u32 NONE_V = 0xffffffffu;

struct DevCx {
    u32 n;
    float thr;
    const float* mat;      // dense: n x n full layout; null for a graph
    const u64* rp;         // sparse: CSR, rows sorted by neighbour id
    const u32* col;
    const float* w;
};

// d(a, b) when (a, b) is an edge of the complex (always, on a dense one)
__device__ __forceinline__ bool edge_w(const DevCx& c, u32 a, u32 b, float& out) {
    if (a == b) {
        out = 0.0f;
        return true;
    }
    if (c.mat) {
        out = c.mat[u64(a) * c.n + b];
        return true;
    }
    u64 lo = c.rp[a], hi = c.rp[a + 1];
    const u64 end = hi;
    while (lo < hi) {
        const u64 mid = lo + ((hi - lo) >> 1);
        if (c.col[mid] < b) lo = mid + 1;
        else hi = mid;
    }
    if (lo < end && c.col[lo] == b) {
        out = c.w[lo];
        return true;
    }
    return false;
}

// C(v, k) from the table B[k * (n + 1) + v]
template <typename I>
__device__ __forceinline__ I binom(const I* B, u32 n, u32 v, u32 k) {
    return B[u64(k) * (n + 1) + v];
}

// vertices of simplex `idx` with k vertices, descending
template <typename I>
__device__ void decode(const I* B, u32 n, I idx, u32 k, u32* vd) {
    u32 top = n;
    for (u32 pos = k; pos >= 1; --pos) {
        const I* row = B + u64(pos) * (n + 1);
        u32 lo = pos - 1, hi = top;   // row[lo] = 0 <= idx
        while (hi - lo > 1) {
            const u32 mid = lo + ((hi - lo) >> 1);
            if (row[mid] <= idx) lo = mid;
            else hi = mid;
        }
        vd[k - pos] = lo;
        idx -= row[lo];
        top = lo;
    }
}

template <typename I>
__device__ __forceinline__ I index_of(const I* B, u32 n, const u32* vd, u32 k) {
    I s = 0;
    for (u32 i = 0; i < k; ++i) s += binom(B, n, vd[i], k - i);
    return s;
}

// U + {w} into out (descending); returns w's position
__device__ __forceinline__ u32 insert_vertex(const u32* U, u32 k, u32 w, u32* out) {
    u32 t = 0;
    while (t < k && U[t] > w) {
        out[t] = U[t];
        ++t;
    }
    out[t] = w;
    for (u32 i = t; i < k; ++i) out[i + 1] = U[i];
    return t;
}

__device__ __forceinline__ bool member(const u32* U, u32 k, u32 w) {
    for (u32 i = 0; i < k; ++i)
        if (U[i] == w) return true;
    return false;
}

// Largest w > above (above = -1: any) outside U (k vertices, descending)
// with every d(u, w) <= d; NONE_V if none.
__device__ u32 zpc_vertex(const DevCx& c, const u32* U, u32 k, float d, long long above) {
    if (c.mat) {
        for (long long wi = (long long)c.n - 1; wi > above; --wi) {
            const u32 w = u32(wi);
            bool ok = true;
            for (u32 i = 0; i < k && ok; ++i) {
                if (U[i] == w) ok = false;
                else ok = c.mat[u64(U[i]) * c.n + w] <= d;
            }
            if (ok) return w;
        }
        return NONE_V;
    }
    // the shortest row drives; the others are probed by binary search
    u32 drv = 0;
    u64 best = ~0ull;
    for (u32 i = 0; i < k; ++i) {
        const u64 deg = c.rp[U[i] + 1] - c.rp[U[i]];
        if (deg < best) {
            best = deg;
            drv = i;
        }
    }
    const u64 lo = c.rp[U[drv]];
    for (u64 p = c.rp[U[drv] + 1]; p > lo; --p) {
        const u32 x = c.col[p - 1];
        if ((long long)x <= above) break;
        if (!(c.w[p - 1] <= d) || member(U, k, x)) continue;
        bool ok = true;
        for (u32 i = 0; i < k && ok; ++i) {
            if (i == drv) continue;
            float wv;
            ok = edge_w(c, U[i], x, wv) && wv <= d;
        }
        if (ok) return x;
    }
    return NONE_V;
}

// Position of the zero-pivot facet of U (k vertices, descending, diameter
// d) — the vertex removed — or k if none (k < 3 has none).
__device__ u32 zpf_position(const DevCx& c, const u32* U, u32 k, float d) {
    u64 common = ~0ull;   // positions on every edge of length d
    bool any = false;
    for (u32 a = 0; a < k; ++a)
        for (u32 b = a + 1; b < k; ++b) {
            float wv;
            if (edge_w(c, U[a], U[b], wv) && wv == d) {
                common &= (u64(1) << a) | (u64(1) << b);
                any = true;
            }
        }
    if (!any) return k;
    const u32 j = u32(__builtin_ctzll(~common));
    return j < k ? j : k;
}

// U without position j into F (k - 1 vertices)
__device__ __forceinline__ void remove_at(const u32* U, u32 k, u32 j, u32* F) {
    for (u32 i = 0, o = 0; i < k; ++i)
        if (i != j) F[o++] = U[i];
}

// Emergent state of a column U (k vertices, diameter d): 0 no cofacet of
// equal diameter; 1 cofacet E = U + {w} with no zero-apparent facet
// (*cof = its index); 2 E has one. `ew` is zpc(U, d) when the caller has it.
template <typename I, u32 KV>
__device__ u8 emergent_state(const DevCx& c, const I* B, const u32* U, u32 k, float d, u32 ew, I* cof) {
    if (ew == NONE_V) return 0;
    u32 E[KV + 2], F[KV + 2];
    insert_vertex(U, k, ew, E);
    *cof = index_of(B, c.n, E, k + 1);
    const u32 j = zpf_position(c, E, k + 1, d);
    if (j > k) return 1;   // no facet of diameter d: impossible for a cofacet of U, kept total
    remove_at(E, k + 1, j, F);
    return zpc_vertex(c, F, k, d, (long long)E[j]) == NONE_V ? 2 : 1;
}

// ---- the next dimension's cofacets ---------------------------------------
// Cofacets of simplex (U, ds) with a new top vertex x > U[0]: every (u, x)
// an edge, cd = max(ds, d(u, x)) <= thr. f(x, cd) per cofacet.
template <u32 KV, typename F>
__device__ void top_cofacets(const DevCx& c, const u32* U, u32 k, float ds, F&& f) {
    if (c.mat) {
        for (u32 x = U[0] + 1; x < c.n; ++x) {
            float cd = ds;
            bool ok = true;
            for (u32 i = 0; i < k && ok; ++i) {
                const float v = c.mat[u64(U[i]) * c.n + x];
                cd = cd < v ? v : cd;
                ok = cd <= c.thr;
            }
            if (ok) f(x, cd);
        }
        return;
    }
    // neighbours of U[0] above it, probed in the other rows
    u64 lo = c.rp[U[0]], hi = c.rp[U[0] + 1];
    const u64 end = hi;
    while (lo < hi) {
        const u64 mid = lo + ((hi - lo) >> 1);
        if (c.col[mid] <= U[0]) lo = mid + 1;
        else hi = mid;
    }
    for (u64 p = lo; p < end; ++p) {
        const u32 x = c.col[p];
        float cd = ds < c.w[p] ? c.w[p] : ds;
        bool ok = cd <= c.thr;
        for (u32 i = 1; i < k && ok; ++i) {
            float wv;
            ok = edge_w(c, U[i], x, wv);
            if (ok) {
                cd = cd < wv ? wv : cd;
                ok = cd <= c.thr;
            }
        }
        if (ok) f(x, cd);
    }
}

template <typename I, u32 KV>
__global__ __launch_bounds__(CB) void cob_count(DevCx c, const I* __restrict__ B, const I* __restrict__ sidx,
                                                const float* __restrict__ sd, u64 count, u32 k,
                                                u32* __restrict__ cnt) {
    for (u64 t = blockIdx.x * u64(CB) + threadIdx.x; t < count; t += u64(gridDim.x) * CB) {
        u32 U[KV];
        decode(B, c.n, sidx[t], k, U);
        u32 m = 0;
        top_cofacets<KV>(c, U, k, sd[t], [&](u32, float) { ++m; });
        cnt[t] = m;
    }
}

template <typename I, u32 KV>
__global__ __launch_bounds__(CB) void cob_emit(DevCx c, const I* __restrict__ B, const I* __restrict__ sidx,
                                               const float* __restrict__ sd, u64 count, u32 k,
                                               const u32* __restrict__ off, I* __restrict__ oidx,
                                               float* __restrict__ od) {
    for (u64 t = blockIdx.x * u64(CB) + threadIdx.x; t < count; t += u64(gridDim.x) * CB) {
        u32 U[KV];
        const I s = sidx[t];
        decode(B, c.n, s, k, U);
        u32 o = off[t];
        top_cofacets<KV>(c, U, k, sd[t], [&](u32 x, float cd) {
            oidx[o] = s + binom(B, c.n, x, k + 1);
            od[o] = cd;
            ++o;
        });
    }
}

template <typename I>
__device__ __forceinline__ bool sorted_has(const I* __restrict__ keys, u64 count, I v) {
    u64 lo = 0, hi = count;
    while (lo < hi) {
        const u64 mid = lo + ((hi - lo) >> 1);
        if (keys[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo < count && keys[lo] == v;
}

// Per cofacet (k vertices): keep[t] = 1 for a column, with its emergent
// state and cofacet; skipped (apparent) counted.
template <typename I, u32 KV>
__global__ __launch_bounds__(CB) void col_flags(DevCx c, const I* __restrict__ B, const I* __restrict__ cidx,
                                                const float* __restrict__ cdd, u32 count, u32 k,
                                                const I* __restrict__ piv, u64 npiv, u32* __restrict__ keep,
                                                u8* __restrict__ state, I* __restrict__ cof,
                                                u32* __restrict__ skipped) {
    u32 sk = 0;
    for (u32 t = blockIdx.x * CB + threadIdx.x; t < count; t += gridDim.x * CB) {
        const I s = cidx[t];
        const float d = cdd[t];
        u32 kp = 0;
        u8 st = 0;
        I cf = 0;
        if (!(npiv && sorted_has(piv, npiv, s))) {
            u32 U[KV + 1], E[KV + 2], F[KV + 1];
            decode(B, c.n, s, k, U);
            const u32 ew = zpc_vertex(c, U, k, d, -1);
            bool apparent = false;
            if (ew != NONE_V) {
                const u32 pw = insert_vertex(U, k, ew, E);
                apparent = zpf_position(c, E, k + 1, d) == pw;
            }
            if (!apparent) {
                const u32 j = zpf_position(c, U, k, d);
                if (j < k) {
                    remove_at(U, k, j, F);
                    apparent = zpc_vertex(c, F, k - 1, d, (long long)U[j]) == NONE_V;
                }
            }
            if (apparent) {
                ++sk;
            } else {
                kp = 1;
                st = emergent_state<I, KV>(c, B, U, k, d, ew, &cf);
            }
        }
        keep[t] = kp;
        state[t] = st;
        cof[t] = cf;
    }
    for (u32 off = 32; off > 0; off >>= 1) sk += __shfl_xor(sk, off);
    if ((threadIdx.x & 63) == 0 && sk) atomicAdd(skipped, sk);
}

// Emergent states of given columns (k vertices)
template <typename I, u32 KV>
__global__ __launch_bounds__(CB) void emergent_kernel(DevCx c, const I* __restrict__ B, const I* __restrict__ idx,
                                                      const float* __restrict__ dd, u64 count, u32 k,
                                                      u8* __restrict__ state, I* __restrict__ cof) {
    for (u64 t = blockIdx.x * u64(CB) + threadIdx.x; t < count; t += u64(gridDim.x) * CB) {
        u32 U[KV + 1];
        decode(B, c.n, idx[t], k, U);
        const float d = dd[t];
        I cf = 0;
        state[t] = emergent_state<I, KV>(c, B, U, k, d, zpc_vertex(c, U, k, d, -1), &cf);
        cof[t] = cf;
    }
}

template <typename I>
__global__ void col_scatter(const u32* __restrict__ keep, const u32* __restrict__ pos, u32 count,
                            const I* __restrict__ cidx, const float* __restrict__ cdd, const u8* __restrict__ st,
                            const I* __restrict__ cof, I* __restrict__ oidx, float* __restrict__ od,
                            u8* __restrict__ ost, I* __restrict__ ocof) {
    for (u32 t = blockIdx.x * blockDim.x + threadIdx.x; t < count; t += gridDim.x * blockDim.x)
        if (keep[t]) {
            const u32 o = pos[t];
            oidx[o] = cidx[t];
            od[o] = cdd[t];
            ost[o] = st[t];
            ocof[o] = cof[t];
        }
}

// sort keys: one 64-bit word of the index, or the diameter descending
template <typename I>
__global__ void key_word(const I* __restrict__ idx, const u32* __restrict__ perm, u32 count, int word,
                         u64* __restrict__ keys) {
    for (u32 t = blockIdx.x * blockDim.x + threadIdx.x; t < count; t += gridDim.x * blockDim.x) {
        const I v = idx[perm ? perm[t] : t];
        keys[t] = word ? u64(v >> 32 >> 32) : u64(v);
    }
}

__global__ void key_desc_d(const float* __restrict__ d, const u32* __restrict__ perm, u32 count,
                           u64* __restrict__ keys) {
    for (u32 t = blockIdx.x * blockDim.x + threadIdx.x; t < count; t += gridDim.x * blockDim.x)
        keys[t] = u64(0x7fffffffu - (__float_as_uint(d[perm[t]]) & 0x7fffffffu));   // -0 as +0
}

// the full n x n layout from the packed lower triangle (i > j at i(i-1)/2 + j)
__global__ void lower_to_full(const float* __restrict__ low, u32 n, float* __restrict__ full) {
    const u64 cells = u64(n) * n;
    for (u64 t = blockIdx.x * u64(blockDim.x) + threadIdx.x; t < cells; t += u64(gridDim.x) * blockDim.x) {
        const u32 i = u32(t / n), j = u32(t % n);
        full[t] = i == j ? 0.0f : (i > j ? low[u64(i) * (i - 1) / 2 + j] : low[u64(j) * (j - 1) / 2 + i]);
    }
}

__global__ void iota_u32(u32* __restrict__ p, u32 count) {
    for (u32 t = blockIdx.x * blockDim.x + threadIdx.x; t < count; t += gridDim.x * blockDim.x) p[t] = t;
}

template <typename I>
__global__ void col_gather(const u32* __restrict__ perm, u32 count, const I* __restrict__ idx,
                           const float* __restrict__ d, const u8* __restrict__ st, const I* __restrict__ cof,
                           I* __restrict__ oidx, float* __restrict__ od, u8* __restrict__ ost,
                           I* __restrict__ ocof) {
    for (u32 t = blockIdx.x * blockDim.x + threadIdx.x; t < count; t += gridDim.x * blockDim.x) {
        const u32 s = perm[t];
        oidx[t] = idx[s];
        od[t] = d[s];
        ost[t] = st[s];
        ocof[t] = cof[s];
    }
}

u32 grid_of(u64 items, u32 block = CB) {
    return u32(std::max<u64>(1, std::min<u64>((items + block - 1) / block, 65536)));
}

template <typename I>
void to_words(const I* src, u64 count, uint64_t* dst, int words) {
    for (u64 t = 0; t < count; ++t) {
        dst[t * words] = u64(src[t]);
        if (words == 2) dst[t * 2 + 1] = u64(src[t] >> 32 >> 32);
    }
}

template <typename I>
std::vector<I> from_words(const uint64_t* src, u64 count, int words) {
    std::vector<I> out(count);
    for (u64 t = 0; t < count; ++t) {
        I v = I(src[t * words]);
        if (words == 2) v |= I(src[t * 2 + 1]) << 32 << 32;
        out[t] = v;
    }
    return out;
}

}  // namespace

// The device side of one complex: its distances, binomial table, the
// simplices assemble starts from, and the index width.
struct Complex {
    u32 n = 0, K = 0;
    float thr = INFINITY;
    int words = 1;
    bool sparse = false;
    DevBuf<float> mat_own;
    const float* mat = nullptr;
    DevBuf<u64> rp;
    DevBuf<u32> col;
    DevBuf<float> w;
    DevBuf<unsigned char> binom;   // (K + 1) x (n + 1) of I
    std::vector<u128> top;         // C(n, k), k = 0..K
    u32 nv = 0;                    // vertices of the current simplices
    u64 count = 0;
    DevBuf<unsigned char> sidx;    // current simplices: I[count]
    DevBuf<float> sd;
    hipStream_t s = nullptr;

    DevCx view() const {
        DevCx c;
        c.n = n;
        c.thr = thr;
        c.mat = sparse ? nullptr : mat;
        c.rp = sparse ? rp.p : nullptr;
        c.col = sparse ? col.p : nullptr;
        c.w = sparse ? w.p : nullptr;
        return c;
    }
};

namespace {

template <typename I>
void upload_binomials(Complex& c) {
    const u64 cols = u64(c.n) + 1;
    std::vector<I> t(cols * (c.K + 1), I(0));
    for (u32 v = 0; v <= c.n; ++v) {
        t[v] = 1;
        for (u32 k = 1; k <= std::min(v, c.K); ++k) {
            I s;
            if (__builtin_add_overflow(t[u64(k - 1) * cols + v - 1], t[u64(k) * cols + v - 1], &s))
                fail(TT_EOVERFLOW, "rips: simplex index overflow");
            t[u64(k) * cols + v] = s;
        }
    }
    // the all-ones index stays free as a sentinel, as the reducer requires
    if (t[u64(c.K) * cols + c.n] == ~I(0)) fail(TT_EOVERFLOW, "rips: simplex index overflow");
    c.top.assign(c.K + 1, 0);
    for (u32 k = 0; k <= c.K; ++k) c.top[k] = u128(t[u64(k) * cols + c.n]);
    c.binom.alloc(t.size() * sizeof(I));
    TTG_HIP(hipMemcpy(c.binom.p, t.data(), t.size() * sizeof(I), hipMemcpyHostToDevice));
}

// Launch f.operator()<KV>() with the smallest vertex capacity that holds
// simplices of k vertices plus two more.
template <typename F>
void with_kv(u32 k, F&& f) {
    if (k + 2 <= 8) f.template operator()<8>();
    else if (k + 2 <= 34) f.template operator()<34>();
    else fail(TT_EINVAL, "rips: maxdim must be <= 30");
}

template <typename I>
void emergent_impl(Complex& c, u32 k, const uint64_t* idx, const float* d, u64 count, uint64_t* cof_out,
                   uint8_t* state_out) {
    if (!count) return;
    Workspace ws;
    const std::vector<I> hidx = from_words<I>(idx, count, c.words);
    I* di = ws.get<I>(count);
    float* dd = ws.get<float>(count);
    u8* dst = ws.get<u8>(count);
    I* dcf = ws.get<I>(count);
    TTG_HIP(hipMemcpyAsync(di, hidx.data(), count * sizeof(I), hipMemcpyHostToDevice, c.s));
    TTG_HIP(hipMemcpyAsync(dd, d, count * sizeof(float), hipMemcpyHostToDevice, c.s));
    const I* B = reinterpret_cast<const I*>(c.binom.p);
    with_kv(k, [&]<u32 KV>() {
        emergent_kernel<I, KV><<<grid_of(count), CB, 0, c.s>>>(c.view(), B, di, dd, count, k, dst, dcf);
    });
    TTG_LAUNCH_CHECK();
    std::vector<I> hcf(count);
    TTG_HIP(hipMemcpyAsync(state_out, dst, count, hipMemcpyDeviceToHost, c.s));
    TTG_HIP(hipMemcpyAsync(hcf.data(), dcf, count * sizeof(I), hipMemcpyDeviceToHost, c.s));
    TTG_HIP(hipStreamSynchronize(c.s));
    to_words(hcf.data(), count, cof_out, c.words);
}

template <typename I>
void seed_impl(Complex& c, u32 k, const uint64_t* idx, const float* d, u64 count) {
    const std::vector<I> hidx = from_words<I>(idx, count, c.words);
    c.sidx.alloc(std::max<u64>(1, count) * sizeof(I));
    c.sd.alloc(std::max<u64>(1, count));
    if (count) {
        TTG_HIP(hipMemcpyAsync(c.sidx.p, hidx.data(), count * sizeof(I), hipMemcpyHostToDevice, c.s));
        TTG_HIP(hipMemcpyAsync(c.sd.p, d, count * sizeof(float), hipMemcpyHostToDevice, c.s));
        TTG_HIP(hipStreamSynchronize(c.s));
    }
    c.nv = k;
    c.count = count;
}

// stable sort of positions [0, m) by the 64-bit keys (bits [0, bits)); the
// permutation carried along
void sort_perm(u64* keys, u32* perm, u32 m, int bits, Workspace& ws, hipStream_t s) {
    if (bits <= 0 || m < 2) return;
    u64* ka = ws.get<u64>(m);
    u32* pa = ws.get<u32>(m);
    radix_sort_u64(keys, perm, m, 0, bits, ka, pa, ws, s);
}

template <typename I>
void assemble_impl(Complex& c, const uint64_t* pivots, u64 npiv, bool keep_next, uint64_t* ncols,
                   uint64_t** out_idx, float** out_d, uint64_t** out_cof, uint8_t** out_state, uint64_t* skipped) {
    const u32 k = c.nv + 1;   // vertices of the cofacets
    if (k > c.K) fail(TT_EINVAL, "tt_complex_assemble: dimension above the complex's maxdim + 1");
    Workspace ws;
    hipStream_t s = c.s;
    const DevCx v = c.view();
    const I* B = reinterpret_cast<const I*>(c.binom.p);
    const I* sidx = reinterpret_cast<const I*>(c.sidx.p);
    // 1. every cofacet with a new top vertex
    u64 total = 0;
    I* cidx = nullptr;
    float* cdd = nullptr;
    if (c.count) {
        u32* cnt = ws.get<u32>(c.count + 1);
        u32* tot = ws.get<u32>(1);
        fill_device_u32(cnt + c.count, 1, 0u, s);
        with_kv(c.nv, [&]<u32 KV>() {
            cob_count<I, KV><<<grid_of(c.count), CB, 0, s>>>(v, B, sidx, c.sd.p, c.count, c.nv, cnt);
        });
        TTG_LAUNCH_CHECK();
        // per-simplex counts fit 32 bits (< n); the total must too
        std::vector<u32> hc(c.count);
        TTG_HIP(hipMemcpyAsync(hc.data(), cnt, c.count * 4, hipMemcpyDeviceToHost, s));
        TTG_HIP(hipStreamSynchronize(s));
        for (u32 x : hc) total += x;
        if (total >= (u64(1) << 32)) fail(TT_ELENGTH, "tt_complex_assemble: more than 2^32 - 1 cofacets");
        exclusive_scan_u32(cnt, cnt, c.count, tot, ws, s);
        cidx = ws.get<I>(std::max<u64>(1, total));
        cdd = ws.get<float>(std::max<u64>(1, total));
        if (total)
            with_kv(c.nv, [&]<u32 KV>() {
                cob_emit<I, KV><<<grid_of(c.count), CB, 0, s>>>(v, B, sidx, c.sd.p, c.count, c.nv, cnt, cidx, cdd);
            });
        TTG_LAUNCH_CHECK();
    }
    const u32 m = u32(total);
    // 2. flags per cofacet
    std::vector<I> hp = from_words<I>(pivots, npiv, c.words);
    std::sort(hp.begin(), hp.end());
    I* dp = ws.get<I>(std::max<u64>(1, npiv));
    if (npiv) TTG_HIP(hipMemcpyAsync(dp, hp.data(), npiv * sizeof(I), hipMemcpyHostToDevice, s));
    u32* keep = ws.get<u32>(u64(m) + 1);
    u8* st = ws.get<u8>(std::max<u32>(1, m));
    I* cf = ws.get<I>(std::max<u32>(1, m));
    u32* dsk = ws.get<u32>(2);
    fill_device_u32(dsk, 2, 0u, s);
    fill_device_u32(keep + m, 1, 0u, s);
    if (m)
        with_kv(k, [&]<u32 KV>() {
            col_flags<I, KV><<<grid_of(m), CB, 0, s>>>(v, B, cidx, cdd, m, k, dp, npiv, keep, st, cf, dsk);
        });
    TTG_LAUNCH_CHECK();
    // 3. compact the columns
    u32* pos = ws.get<u32>(u64(m) + 1);
    exclusive_scan_u32(keep, pos, u64(m) + 1, nullptr, ws, s);
    u32 hm[2] = {0, 0};
    TTG_HIP(hipMemcpyAsync(&hm[0], pos + m, 4, hipMemcpyDeviceToHost, s));
    TTG_HIP(hipMemcpyAsync(&hm[1], dsk, 4, hipMemcpyDeviceToHost, s));
    TTG_HIP(hipStreamSynchronize(s));
    const u32 nc = hm[0];
    *skipped = hm[1];
    I* kidx = ws.get<I>(std::max<u32>(1, nc));
    float* kd = ws.get<float>(std::max<u32>(1, nc));
    u8* kst = ws.get<u8>(std::max<u32>(1, nc));
    I* kcf = ws.get<I>(std::max<u32>(1, nc));
    if (m) col_scatter<I><<<grid_of(m), CB, 0, s>>>(keep, pos, m, cidx, cdd, st, cf, kidx, kd, kst, kcf);
    TTG_LAUNCH_CHECK();
    // 4. sort: index ascending (low word, then high word), then d descending
    u32* perm = ws.get<u32>(std::max<u32>(1, nc));
    u64* keys = ws.get<u64>(std::max<u32>(1, nc));
    std::vector<I> hidx(nc), hcf(nc);
    std::vector<u8> hst(nc);
    std::vector<float> hd(nc);
    if (nc) {
        iota_u32<<<grid_of(nc), CB, 0, s>>>(perm, nc);
        // indices are below C(n, k): only the bits that vary are sorted
        const u128 top = c.top[k];
        const int lo_bits = (top >> 64) ? 64 : bit_length(u64(top));
        const int hi_bits = bit_length(u64(top >> 64));
        key_word<I><<<grid_of(nc), CB, 0, s>>>(kidx, nullptr, nc, 0, keys);
        sort_perm(keys, perm, nc, lo_bits, ws, s);
        if (c.words == 2 && hi_bits) {
            key_word<I><<<grid_of(nc), CB, 0, s>>>(kidx, perm, nc, 1, keys);
            sort_perm(keys, perm, nc, hi_bits, ws, s);
        }
        key_desc_d<<<grid_of(nc), CB, 0, s>>>(kd, perm, nc, keys);
        sort_perm(keys, perm, nc, 31, ws, s);
        I* gi = ws.get<I>(nc);
        float* gd = ws.get<float>(nc);
        u8* gs = ws.get<u8>(nc);
        I* gc = ws.get<I>(nc);
        col_gather<I><<<grid_of(nc), CB, 0, s>>>(perm, nc, kidx, kd, kst, kcf, gi, gd, gs, gc);
        TTG_LAUNCH_CHECK();
        TTG_HIP(hipMemcpyAsync(hidx.data(), gi, u64(nc) * sizeof(I), hipMemcpyDeviceToHost, s));
        TTG_HIP(hipMemcpyAsync(hd.data(), gd, u64(nc) * 4, hipMemcpyDeviceToHost, s));
        TTG_HIP(hipMemcpyAsync(hst.data(), gs, nc, hipMemcpyDeviceToHost, s));
        TTG_HIP(hipMemcpyAsync(hcf.data(), gc, u64(nc) * sizeof(I), hipMemcpyDeviceToHost, s));
    }
    // 5. the cofacets become the next simplices
    if (keep_next) {
        DevBuf<unsigned char> ni;
        DevBuf<float> nd;
        ni.alloc(std::max<u64>(1, total) * sizeof(I));
        nd.alloc(std::max<u64>(1, total));
        if (total) {
            TTG_HIP(hipMemcpyAsync(ni.p, cidx, total * sizeof(I), hipMemcpyDeviceToDevice, s));
            TTG_HIP(hipMemcpyAsync(nd.p, cdd, total * 4, hipMemcpyDeviceToDevice, s));
        }
        TTG_HIP(hipStreamSynchronize(s));
        std::swap(c.sidx.p, ni.p);
        std::swap(c.sidx.n, ni.n);
        std::swap(c.sd.p, nd.p);
        std::swap(c.sd.n, nd.n);
        c.nv = k;
        c.count = total;
    }
    TTG_HIP(hipStreamSynchronize(s));
    *ncols = nc;
    const int wd = c.words;
    *out_idx = static_cast<uint64_t*>(std::malloc(std::max<u64>(1, u64(nc) * wd) * 8));
    *out_cof = static_cast<uint64_t*>(std::malloc(std::max<u64>(1, u64(nc) * wd) * 8));
    *out_d = static_cast<float*>(std::malloc(std::max<u64>(1, nc) * 4));
    *out_state = static_cast<uint8_t*>(std::malloc(std::max<u64>(1, nc)));
    if (!*out_idx || !*out_cof || !*out_d || !*out_state) fail(TT_EDEVICE, "out of host memory");
    to_words(hidx.data(), nc, *out_idx, wd);
    to_words(hcf.data(), nc, *out_cof, wd);
    if (nc) {
        std::memcpy(*out_d, hd.data(), u64(nc) * 4);
        std::memcpy(*out_state, hst.data(), nc);
    }
}

}  // namespace

}  // namespace ttg

using namespace ttg;

struct tt_complex {
    Complex c;
};

namespace {

void check_words(int words) {
    if (words != 1 && words != 2) fail(TT_EINVAL, "tt_complex: index_words must be 1 or 2");
}

void finish_create(tt_complex* h, uint32_t maxdim, int words) {
    Complex& c = h->c;
    if (maxdim + 2 > 32) fail(TT_EINVAL, "rips: maxdim must be <= 30");
    c.K = maxdim + 2;
    c.words = words;
    if (words == 2) upload_binomials<u128>(c);
    else upload_binomials<u64>(c);
}

}  // namespace

extern "C" {

tt_status tt_complex_dense(const float* mat, uint32_t n, int layout, int on_device, float threshold,
                           uint32_t maxdim, int index_words, tt_complex** out) {
    return guarded([&] {
        if (!out) fail(TT_EINVAL, "tt_complex_dense: null output");
        *out = nullptr;
        check_words(index_words);
        if (n && !mat) fail(TT_EINVAL, "tt_complex_dense: null matrix");
        if (layout != TT_LAYOUT_FULL && layout != TT_LAYOUT_LOWER) fail(TT_EINVAL, "tt_complex_dense: bad layout");
        if (on_device && layout != TT_LAYOUT_FULL) fail(TT_EINVAL, "tt_complex_dense: a device matrix must be full");
        if (std::isnan(threshold)) fail(TT_EINVAL, "persistence: threshold must be non-negative or +inf");
        auto h = std::make_unique<tt_complex>();
        Complex& c = h->c;
        c.n = n;
        c.thr = threshold;
        c.sparse = false;
        if (on_device) {
            c.mat = mat;
        } else {
            c.mat_own.alloc(std::max<u64>(1, u64(n) * n));
            if (n && layout == TT_LAYOUT_FULL) {
                TTG_HIP(hipMemcpy(c.mat_own.p, mat, u64(n) * n * 4, hipMemcpyHostToDevice));
            } else if (n) {
                const u64 packed = u64(n) * (n - 1) / 2;
                DevBuf<float> low(std::max<u64>(1, packed));
                if (packed) TTG_HIP(hipMemcpy(low.p, mat, packed * 4, hipMemcpyHostToDevice));
                lower_to_full<<<grid_of(u64(n) * n), CB, 0, nullptr>>>(low.p, n, c.mat_own.p);
                TTG_LAUNCH_CHECK();
                TTG_HIP(hipDeviceSynchronize());
            }
            c.mat = c.mat_own.p;
        }
        finish_create(h.get(), maxdim, index_words);
        *out = h.release();
    });
}

tt_status tt_complex_sparse(uint32_t n, const uint64_t* row_ptr, const uint32_t* col, const float* w,
                            float threshold, uint32_t maxdim, int index_words, tt_complex** out) {
    return guarded([&] {
        if (!out) fail(TT_EINVAL, "tt_complex_sparse: null output");
        *out = nullptr;
        check_words(index_words);
        if (!row_ptr) fail(TT_EINVAL, "tt_complex_sparse: null row_ptr");
        if (std::isnan(threshold)) fail(TT_EINVAL, "persistence: threshold must be non-negative or +inf");
        const u64 nnz = row_ptr[n];
        if (nnz && (!col || !w)) fail(TT_EINVAL, "tt_complex_sparse: null columns");
        for (u32 v = 0; v < n; ++v) {
            if (row_ptr[v + 1] < row_ptr[v]) fail(TT_EINVAL, "sparse: row_ptr must be non-decreasing");
            for (u64 t = row_ptr[v]; t < row_ptr[v + 1]; ++t) {
                if (col[t] >= n) fail(TT_EINVAL, "sparse: column index out of range");
                if (t > row_ptr[v] && col[t] <= col[t - 1]) fail(TT_EINVAL, "sparse: rows must be sorted and unique");
            }
        }
        auto h = std::make_unique<tt_complex>();
        Complex& c = h->c;
        c.n = n;
        c.thr = threshold;
        c.sparse = true;
        c.rp.alloc(u64(n) + 1);
        c.col.alloc(std::max<u64>(1, nnz));
        c.w.alloc(std::max<u64>(1, nnz));
        TTG_HIP(hipMemcpy(c.rp.p, row_ptr, (u64(n) + 1) * 8, hipMemcpyHostToDevice));
        if (nnz) {
            TTG_HIP(hipMemcpy(c.col.p, col, nnz * 4, hipMemcpyHostToDevice));
            TTG_HIP(hipMemcpy(c.w.p, w, nnz * 4, hipMemcpyHostToDevice));
        }
        finish_create(h.get(), maxdim, index_words);
        *out = h.release();
    });
}

tt_status tt_complex_emergent(tt_complex* h, uint32_t nv, const uint64_t* idx, const float* d, uint64_t count,
                              uint64_t* cof, uint8_t* state) {
    return guarded([&] {
        if (!h) fail(TT_EINVAL, "tt_complex_emergent: null complex");
        if (count && (!idx || !d || !cof || !state)) fail(TT_EINVAL, "tt_complex_emergent: null buffer");
        if (nv < 1 || nv + 1 > h->c.K) fail(TT_EINVAL, "tt_complex_emergent: dimension above the complex's maxdim");
        if (h->c.words == 2) emergent_impl<u128>(h->c, nv, idx, d, count, cof, state);
        else emergent_impl<u64>(h->c, nv, idx, d, count, cof, state);
    });
}

tt_status tt_complex_seed(tt_complex* h, uint32_t nv, const uint64_t* idx, const float* d, uint64_t count) {
    return guarded([&] {
        if (!h) fail(TT_EINVAL, "tt_complex_seed: null complex");
        if (count && (!idx || !d)) fail(TT_EINVAL, "tt_complex_seed: null buffer");
        if (nv < 1 || nv + 1 > h->c.K) fail(TT_EINVAL, "tt_complex_seed: dimension above the complex's maxdim");
        if (h->c.words == 2) seed_impl<u128>(h->c, nv, idx, d, count);
        else seed_impl<u64>(h->c, nv, idx, d, count);
    });
}

tt_status tt_complex_assemble(tt_complex* h, const uint64_t* pivots, uint64_t npivots, int keep_next,
                              uint64_t* ncols, uint64_t** idx, float** d, uint64_t** cof, uint8_t** state,
                              uint64_t* skipped) {
    return guarded([&] {
        if (!h || !ncols || !idx || !d || !cof || !state || !skipped)
            fail(TT_EINVAL, "tt_complex_assemble: null argument");
        if (npivots && !pivots) fail(TT_EINVAL, "tt_complex_assemble: null pivots");
        *ncols = 0;
        *idx = *cof = nullptr;
        *d = nullptr;
        *state = nullptr;
        if (!h->c.nv) fail(TT_EINVAL, "tt_complex_assemble: no simplices seeded");
        if (h->c.words == 2)
            assemble_impl<u128>(h->c, pivots, npivots, keep_next != 0, ncols, idx, d, cof, state, skipped);
        else
            assemble_impl<u64>(h->c, pivots, npivots, keep_next != 0, ncols, idx, d, cof, state, skipped);
    });
}

void tt_complex_destroy(tt_complex* h) { delete h; }

}  // extern "C"
