// SPDX-License-Identifier: MIT
// reducer.hpp — host-side persistent cohomology of a Rips filtration from
// front-end outputs computed on the GPU.
//
// The algorithm is Ripser's (U. Bauer, "Ripser: efficient computation of
// Vietoris-Rips persistence barcodes", J. Appl. Comput. Topol. 2021), the one
// the reference's coboundary engine implements (include/turbotops/rips.hpp:
// 489-1032): simplices are addressed by the combinatorial number system;
// columns are coboundaries enumerated on demand and reduced in a binary heap
// with mod-2 cancellation; apparent pairs and emergent pairs skip reduction;
// clearing drops every column that is already a pivot of the previous
// dimension. Filtration order is (diameter ascending, index descending)
// (rips.hpp:116-118). The diagram does not depend on the reduction strategy,
// only on the filtration.
//
// Dimension 0 does not run here: the GPU hands over the kept edges in
// filtration order with their apparent flags (edges.hip), and the union-find
// below only splits them into H0 deaths and dimension-1 columns.
//
// Besides the bars, every finite pair records its birth and death simplices
// (for the gradient of a diagram with respect to the points).
#pragma once

#include <algorithm>
#include <array>
#include <bit>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

namespace tops::cob {

using u8 = uint8_t;
using u32 = uint32_t;
using u64 = uint64_t;
using u128 = unsigned __int128;
using f32 = float;

template <class I>
constexpr I kNone = ~I(0);

// Bars per dimension in discovery order; birth/death simplex indices of each
// bar (death simplex kNone for essential classes, birth = vertex for H0).
struct Diagram {
    std::vector<std::vector<f32>> birth, death;
    std::vector<std::vector<u128>> birth_simplex, death_simplex;
    void ensure(u32 d) {
        if (birth.size() <= d) {
            birth.resize(d + 1);
            death.resize(d + 1);
            birth_simplex.resize(d + 1);
            death_simplex.resize(d + 1);
        }
    }
    void add(u32 dim, f32 b, f32 e, u128 bs, u128 ds) {
        ensure(dim);
        birth[dim].push_back(b);
        death[dim].push_back(e);
        birth_simplex[dim].push_back(bs);
        death_simplex[dim].push_back(ds);
    }
};

struct Telemetry {
    bool wide_index = false;
    u64 edges = 0;
    std::vector<u64> columns, apparent, emergent;
};

// C(i, k) for i <= n, k <= K with overflow detection.
template <class I>
struct Binomials {
    u32 n = 0, K = 0;
    std::vector<I> t;   // t[k * (n + 1) + i]
    Binomials(u32 nn, u32 KK) : n(nn), K(KK), t(u64(nn + 1) * (KK + 1), I(0)) {
        for (u32 i = 0; i <= n; ++i) {
            t[i] = 1;
            for (u32 k = 1; k <= std::min(i, K); ++k) {
                I s;
                if (__builtin_add_overflow(at(i - 1, k - 1), at(i - 1, k), &s))
                    throw std::overflow_error("rips: simplex index overflow");
                t[u64(k) * (n + 1) + i] = s;
            }
        }
    }
    I at(u32 i, u32 k) const { return k > K ? I(0) : t[u64(k) * (n + 1) + i]; }
    I operator()(u32 i, u32 k) const { return at(i, k); }
};

template <class I>
bool index_fits(u32 n, u32 maxdim) {
    try {
        Binomials<I> b(n, maxdim + 2);
        return b(n, maxdim + 2) < kNone<I>;
    } catch (const std::overflow_error&) {
        return false;
    }
}

// ---- complexes ----------------------------------------------------------------
// Both expose n, threshold, dist(i, j) and
//   cofacets(vd, k, diam, all, cap, emit): vertices w not in the simplex (vd:
//   k vertices, descending) with max(diam, d(v, w) for v in vd) <= cap, in
//   descending w (all == false: only w > vd[0]); emit(w, t, cd) with t the
//   number of simplex vertices above w; emit returns false to stop.

// Dense distances, full (n x n) or lower (i(i-1)/2 + j) layout.
struct DenseView {
    u32 n = 0;
    f32 threshold = INFINITY;
    bool full = true;
    const f32* mat = nullptr;

    f32 dist(u32 i, u32 j) const {
        if (i == j) return 0.0f;
        if (full) return mat[u64(i) * n + j];
        if (i < j) std::swap(i, j);
        return mat[u64(i) * (i - 1) / 2 + j];
    }

    template <class Emit>
    void cofacets(const u32* vd, u32 k, f32 diam, bool all, f32 cap, Emit&& emit) const {
        const u32 stop = all ? 0 : vd[0] + 1;
        u32 t = 0;
        if (full) {
            constexpr u32 B = 64;
            f32 mx[B];
            for (u32 hi = n; hi > stop;) {
                const u32 lo = hi - std::min<u32>(B, hi - stop);
                const u32 len = hi - lo;
                const f32* r0 = mat + u64(vd[0]) * n + lo;
                for (u32 j = 0; j < len; ++j) mx[j] = r0[j];
                for (u32 i = 1; i < k; ++i) {
                    const f32* ri = mat + u64(vd[i]) * n + lo;
                    for (u32 j = 0; j < len; ++j) mx[j] = mx[j] < ri[j] ? ri[j] : mx[j];
                }
                for (u32 j = len; j-- > 0;) {
                    const f32 m = mx[j] < diam ? diam : mx[j];
                    if (!(m <= cap)) continue;
                    const u32 w = lo + j;
                    while (t < k && vd[t] > w) ++t;
                    if (t < k && vd[t] == w) continue;
                    if (!emit(w, t, m)) return;
                }
                hi = lo;
            }
            return;
        }
        for (u32 wi = n; wi-- > stop;) {
            const u32 w = wi;
            while (t < k && vd[t] > w) ++t;
            if (t < k && vd[t] == w) continue;
            f32 cd = diam;
            bool ok = true;
            for (u32 i = 0; i < k; ++i) {
                const f32 dv = dist(vd[i], w);
                if (dv > cd) {
                    cd = dv;
                    if (!(cd <= cap)) {
                        ok = false;
                        break;
                    }
                }
            }
            if (ok && !(cd <= cap)) ok = false;
            if (ok && !emit(w, t, cd)) return;
        }
    }
};

// A weighted graph in CSR form (rows sorted by neighbour id, both
// directions, no loops): cofacets are common neighbours.
struct SparseView {
    u32 n = 0;
    f32 threshold = INFINITY;
    const u64* row_ptr = nullptr;
    const u32* col = nullptr;
    const f32* w = nullptr;

    f32 dist(u32 i, u32 j) const {
        if (i == j) return 0.0f;
        const u32* b = col + row_ptr[i];
        const u32* e = col + row_ptr[i + 1];
        const u32* it = std::lower_bound(b, e, j);
        return (it != e && *it == j) ? w[it - col] : INFINITY;
    }

    template <class Emit>
    void cofacets(const u32* vd, u32 k, f32 diam, bool all, f32 cap, Emit&& emit) const {
        constexpr u32 KM = 33;
        u64 p[KM], lo[KM];
        for (u32 i = 0; i < k; ++i) {
            lo[i] = row_ptr[vd[i]];
            const u64 e = row_ptr[vd[i] + 1];
            if (e == lo[i]) return;
            p[i] = e - 1;
        }
        const u32 stop = all ? 0 : vd[0] + 1;
        for (;;) {
            // the largest id present in every row at or below the cursors
            u32 x = col[p[0]];
            bool agree = false;
            while (!agree) {
                agree = true;
                for (u32 i = 0; i < k; ++i) {
                    while (col[p[i]] > x) {
                        if (p[i] == lo[i]) return;
                        --p[i];
                    }
                    if (col[p[i]] < x) {
                        x = col[p[i]];
                        agree = false;
                    }
                }
            }
            if (x < stop) return;
            u32 t = 0;
            while (t < k && vd[t] > x) ++t;
            f32 cd = diam;
            for (u32 i = 0; i < k; ++i) cd = cd < w[p[i]] ? w[p[i]] : cd;
            if (cd <= cap && !emit(x, t, cd)) return;
            if (p[0] == lo[0]) return;
            --p[0];
        }
    }
};

// ---- pivot map (open addressing) -------------------------------------------
template <class I>
struct PivotTable {
    std::vector<I> key;
    std::vector<u64> val;
    u64 mask = 0, used = 0;
    void reset(u64 expected) {
        u64 c = 64;
        while (c < 2 * expected + 16) c <<= 1;
        key.assign(c, kNone<I>);
        val.assign(c, 0);
        mask = c - 1;
        used = 0;
    }
    static u64 mix(I k) {
        u64 x = u64(k) ^ (u64(k >> 32 >> 32) * 0x9E3779B97F4A7C15ull);
        x ^= x >> 31;
        x *= 0xBF58476D1CE4E5B9ull;
        x ^= x >> 29;
        x *= 0x94D049BB133111EBull;
        return x ^ (x >> 32);
    }
    u64 find(I k) const {
        for (u64 s = mix(k) & mask;; s = (s + 1) & mask) {
            if (key[s] == k) return val[s];
            if (key[s] == kNone<I>) return ~0ull;
        }
    }
    bool has(I k) const { return find(k) != ~0ull; }
    void put(I k, u64 v) {
        if (2 * (used + 1) > key.size()) grow();
        u64 s = mix(k) & mask;
        while (key[s] != kNone<I> && key[s] != k) s = (s + 1) & mask;
        if (key[s] == kNone<I>) ++used;
        key[s] = k;
        val[s] = v;
    }
    void grow() {
        std::vector<I> ok;
        std::vector<u64> ov;
        ok.swap(key);
        ov.swap(val);
        reset(ok.size());
        for (u64 i = 0; i < ok.size(); ++i)
            if (ok[i] != kNone<I>) put(ok[i], ov[i]);
    }
};

// ---- monotone priority queue ---------------------------------------------------
// A radix heap (Ahuja, Mehlhorn, Orlin and Tarjan 1990) over 128-bit keys:
// valid when no key pushed is below the last key popped, which holds for a
// merge of sorted streams. Bucket b holds the items whose key first differs
// from `last` in bit b - 1 (bucket 0: equal to `last`); when bucket 0 runs
// dry, the lowest non-empty bucket is split around its minimum. Each item
// moves down at most 128 buckets over its life, against O(log size) cache
// misses per operation for a binary heap of tens of thousands of streams.
class RadixHeap {
  public:
    void clear() {
        for (int i = 0; i < 129; ++i)
            if (i == 0 || ((occ_[(i - 1) >> 6] >> ((i - 1) & 63)) & 1)) b_[i].clear();
        occ_[0] = occ_[1] = 0;
        n_ = 0;
        last_ = 0;
    }
    bool empty() const { return n_ == 0; }
    u128 last() const { return last_; }
    void push(u128 key, u32 v) {
        put(Item{key, v});
        ++n_;
    }
    // the minimum key (empty() must be false)
    u128 top_key() {
        refill();
        return last_;
    }
    // pops an item with the minimum key, which becomes `last`
    u32 pop() {
        refill();
        const u32 v = b_[0].back().v;
        b_[0].pop_back();
        --n_;
        return v;
    }
    // true while items equal to the last key popped remain
    bool has_last() const { return !b_[0].empty(); }
    u32 pop_last() {
        const u32 v = b_[0].back().v;
        b_[0].pop_back();
        --n_;
        return v;
    }

  private:
    struct Item {
        u128 key;
        u32 v;
    };
    std::array<std::vector<Item>, 129> b_;
    u64 occ_[2] = {0, 0};   // bit b - 1: bucket b (1..128) is non-empty
    size_t n_ = 0;
    u128 last_ = 0;
    void put(const Item& e) {
        const int b = slot(e.key);
        b_[b].push_back(e);
        if (b) occ_[(b - 1) >> 6] |= u64(1) << ((b - 1) & 63);
    }
    int slot(u128 key) const {
        const u128 x = key ^ last_;
        if (!x) return 0;
        const u64 hi = u64(x >> 64);
        return hi ? 128 - __builtin_clzll(hi) : 64 - __builtin_clzll(u64(x));
    }
    void refill() {
        if (!b_[0].empty()) return;
        const int i = occ_[0] ? 1 + __builtin_ctzll(occ_[0]) : 65 + __builtin_ctzll(occ_[1]);
        std::vector<Item>& src = b_[i];
        u128 mn = src[0].key;
        for (const Item& e : src) mn = e.key < mn ? e.key : mn;
        last_ = mn;
        occ_[(i - 1) >> 6] &= ~(u64(1) << ((i - 1) & 63));
        for (const Item& e : src) put(e);   // all land below i
        src.clear();
    }
};

// ---- the reducer ---------------------------------------------------------------
template <class I, class Cx>
class Cohomology {
  public:
    // A simplex in the filtration. Packed: 12 bytes with 64-bit indices, 20
    // with 128-bit ones (32 unpacked), which the column heaps move around.
    struct __attribute__((packed)) S {
        I idx;
        f32 d;
    };
    // a precedes b in the filtration
    static bool before(const S& a, const S& b) { return a.d < b.d || (a.d == b.d && a.idx > b.idx); }

    Cohomology(const Cx& cx, u32 maxdim, Diagram& out, Telemetry& tel)
        : cx_(cx), n_(cx.n), maxdim_(maxdim), B_(cx.n, maxdim + 2), out_(out), tel_(tel) {}

    // host threads for assemble() (the columns of the next dimension)
    void set_threads(u32 t) { threads_ = std::max<u32>(1, t); }

    // Dimension 0 from the GPU's edges (filtration order) and apparent flags.
    void dim0(const u64* idx, const f32* d, const u8* apparent, u64 m) {
        std::vector<u32> parent(n_), rank(n_, 0);
        for (u32 v = 0; v < n_; ++v) parent[v] = v;
        auto find = [&](u32 x) {
            while (parent[x] != x) {
                parent[x] = parent[parent[x]];
                x = parent[x];
            }
            return x;
        };
        cols_.clear();
        u64 skipped = 0;
        for (u64 e = 0; e < m; ++e) {
            u32 a, b;
            decode_edge(idx[e], a, b);
            u32 ra = find(a), rb = find(b);
            if (ra != rb) {
                if (rank[ra] < rank[rb]) std::swap(ra, rb);
                parent[rb] = ra;
                if (rank[ra] == rank[rb]) ++rank[ra];
                if (d[e] > 0.0f) out_.add(0, 0.0f, d[e], u128(a), u128(idx[e]));
            } else if (maxdim_ >= 1) {
                if (apparent && apparent[e]) ++skipped;
                else cols_.push_back(S{I(idx[e]), d[e]});
            }
        }
        for (u32 v = 0; v < n_; ++v)
            if (find(v) == v) out_.add(0, 0.0f, INFINITY, u128(v), u128(kNone<I>));
        std::reverse(cols_.begin(), cols_.end());
        if (maxdim_ >= 2) {
            simplices_.resize(m);
            for (u64 e = 0; e < m; ++e) simplices_[e] = S{I(idx[e]), d[e]};
        }
        tel_.edges = m;
        tel_.apparent.push_back(skipped);
    }

    // dim-0 apparent test of one edge (rips.hpp:844-849): the edge is the
    // zero-pivot facet of its zero-pivot cofacet.
    bool edge_apparent(u64 idx, f32 d) const {
        u32 vd[2];
        decode_edge(idx, vd[0], vd[1]);
        return apparent_cofacet_of(S{I(idx), d}, vd, 2).idx != kNone<I>;
    }

    void run_from_dim1() {
        for (u32 dim = 1; dim <= maxdim_; ++dim) {
            if (!(dim == 1 && reduce_dim1_lazy())) reduce(dim);
            if (dim < maxdim_) assemble(dim);
        }
    }

    // ---- index arithmetic -------------------------------------------------
    // largest v < top with C(v, k) <= idx
    u32 max_vertex(I idx, u32 k, u32 top) const {
        u32 lo = k - 1, hi = top;   // C(lo, k) = 0 <= idx
        while (hi - lo > 1) {
            const u32 mid = lo + (hi - lo) / 2;
            if (B_(mid, k) <= idx) lo = mid;
            else hi = mid;
        }
        return lo;
    }
    void decode(I idx, u32 nv, u32* vd) const {
        u32 top = n_;
        for (u32 i = 0; i < nv; ++i) {
            const u32 k = nv - i;
            const u32 v = max_vertex(idx, k, top);
            vd[i] = v;
            idx -= B_(v, k);
            top = v;
        }
    }
    static void decode_edge(u64 idx, u32& a, u32& b) {
        u64 v = u64((1.0 + std::sqrt(1.0 + 8.0 * double(idx))) * 0.5);
        while (v * (v - 1) / 2 > idx) --v;
        while ((v + 1) * v / 2 <= idx) ++v;
        a = u32(v);
        b = u32(idx - v * (v - 1) / 2);
    }
    I index_of(const u32* vd, u32 nv) const {
        I s = 0;
        for (u32 i = 0; i < nv; ++i) s += B_(vd[i], nv - i);
        return s;
    }

  private:
    static constexpr u32 KM = 33;

    // cofacets of (vd, nv, diam) as simplices, in descending index order
    template <class F>
    void cofacet_simplices(const u32* vd, u32 nv, f32 diam, bool all, f32 cap, F&& f) const {
        I pre[KM + 1], suf[KM + 1];   // pre[t] = sum_{i<t} C(vd[i], nv+1-i); suf[t] = sum_{i>=t} C(vd[i], nv-i)
        pre[0] = 0;
        for (u32 i = 0; i < nv; ++i) pre[i + 1] = pre[i] + B_(vd[i], nv + 1 - i);
        suf[nv] = 0;
        for (u32 i = nv; i-- > 0;) suf[i] = suf[i + 1] + B_(vd[i], nv - i);
        cx_.cofacets(vd, nv, diam, all, cap,
                     [&](u32 w, u32 t, f32 cd) { return f(S{pre[t] + B_(w, nv + 1 - t) + suf[t], cd}, w, t); });
    }

    // zero pivot cofacet: the first cofacet (descending index) of diameter d
    S zero_pivot_cofacet(const u32* vd, u32 nv, f32 d, u32* cv) const {
        S found{kNone<I>, 0.0f};
        u32 fw = 0, ft = 0;
        cofacet_simplices(vd, nv, d, true, d, [&](const S& c, u32 w, u32 t) {
            found = c;
            fw = w;
            ft = t;
            return false;
        });
        if (found.idx != kNone<I>) {
            for (u32 i = 0; i < ft; ++i) cv[i] = vd[i];
            cv[ft] = fw;
            for (u32 i = ft; i < nv; ++i) cv[i + 1] = vd[i];
        }
        return found;
    }

    // zero pivot facet: the first facet (ascending index) of diameter d
    S zero_pivot_facet(const u32* vd, u32 nv, f32 d, u32* fv) const {
        if (nv < 2) return S{kNone<I>, 0.0f};
        for (u32 j = 0; j < nv; ++j) {
            bool hit = false;
            for (u32 a = 0; a < nv && !hit; ++a) {
                if (a == j) continue;
                for (u32 b = a + 1; b < nv && !hit; ++b)
                    if (b != j && cx_.dist(vd[a], vd[b]) == d) hit = true;
            }
            if (hit) {
                u32 o = 0;
                for (u32 i = 0; i < nv; ++i)
                    if (i != j) fv[o++] = vd[i];
                return S{index_of(fv, nv - 1), d};
            }
        }
        return S{kNone<I>, 0.0f};
    }

    // (f, s) is an apparent pair with s the cofacet: returns f, else none
    S apparent_facet_of(const S& s, const u32* vd, u32 nv) const {
        u32 fv[KM], cv[KM + 1];
        const S f = zero_pivot_facet(vd, nv, s.d, fv);
        if (f.idx == kNone<I>) return f;
        return zero_pivot_cofacet(fv, nv - 1, f.d, cv).idx == s.idx ? f : S{kNone<I>, 0.0f};
    }
    // (s, c) is an apparent pair with s the facet: returns c, else none
    S apparent_cofacet_of(const S& s, const u32* vd, u32 nv) const {
        u32 fv[KM + 1], cv[KM + 1];
        const S c = zero_pivot_cofacet(vd, nv, s.d, cv);
        if (c.idx == kNone<I>) return c;
        return zero_pivot_facet(cv, nv + 1, c.d, fv).idx == s.idx ? c : S{kNone<I>, 0.0f};
    }
    S apparent_facet_of(const S& s, u32 nv) const {
        u32 vd[KM];
        decode(s.idx, nv, vd);
        return apparent_facet_of(s, vd, nv);
    }

    // ---- working column: binary min-heap in filtration order -------------
    struct Heap {
        std::vector<S> h;
        static bool later(const S& a, const S& b) { return before(b, a); }
        void clear() { h.clear(); }
        void push(const S& e) {
            h.push_back(e);
            std::push_heap(h.begin(), h.end(), later);
        }
        void load(std::vector<S>& v) {
            h.swap(v);
            std::make_heap(h.begin(), h.end(), later);
        }
        S pop_top() {
            std::pop_heap(h.begin(), h.end(), later);
            const S e = h.back();
            h.pop_back();
            return e;
        }
        // earliest surviving entry after mod-2 cancellation, left in the heap
        S pivot() {
            while (!h.empty()) {
                const S p = pop_top();
                if (!h.empty() && h.front().idx == p.idx) {
                    pop_top();
                    continue;
                }
                push(p);
                return p;
            }
            return S{kNone<I>, 0.0f};
        }
    };

    // The working coboundary of reduce(): a binary heap, or, where every
    // cofacet index fits 64 bits (so a simplex is one 128-bit order key), a
    // radix heap. Its pivot only moves later in the filtration: an added
    // column's coboundary has the current pivot as its own, so its entries
    // before the pivot cancel among themselves and are not pushed.
    bool radix_ = false;
    RadixHeap col_;

    void col_push(const S& c) {
        if (!radix_) {
            heap_.push(c);
            return;
        }
        const u128 k = order_key(c);
        if (k >= col_.last()) col_.push(k, 0);
    }
    void col_load(std::vector<S>& v) {
        if (!radix_) {
            heap_.load(v);
            return;
        }
        col_.clear();
        for (const S& c : v) col_.push(order_key(c), 0);
    }
    // the whole working column in filtration order after mod-2 cancellation,
    // taken out of it
    template <class F>
    void col_drain(F&& f) {
        if (!radix_) {
            for (S p = heap_.pivot(); p.idx != kNone<I>; p = heap_.pivot()) {
                heap_.pop_top();
                f(p);
            }
            return;
        }
        while (!col_.empty()) {
            const u128 k = col_.top_key();
            u32 copies = 0;
            while (col_.has_last()) {
                col_.pop_last();
                ++copies;
            }
            if (copies & 1) f(key_simplex(k));
        }
    }
    // earliest surviving entry after mod-2 cancellation, left in the column
    S col_pivot() {
        if (!radix_) return heap_.pivot();
        while (!col_.empty()) {
            const u128 k = col_.top_key();
            u32 copies = 0;
            while (col_.has_last()) {
                col_.pop_last();
                ++copies;
            }
            if (copies & 1) {
                col_.push(k, 0);
                return key_simplex(k);
            }
        }
        return S{kNone<I>, 0.0f};
    }

    void push_coboundary(const S& s, u32 nv) {
        u32 vd[KM];
        decode(s.idx, nv, vd);
        cofacet_simplices(vd, nv, s.d, true, cx_.threshold, [&](const S& c, u32, u32) {
            col_push(c);
            return true;
        });
    }


    // ---- dimension 1 on a dense full matrix: lazy cofacet streams --------
    // The same column reduction as reduce(1) — same columns, same pivots,
    // same pairs — with the working coboundary held as a merge of sorted
    // streams, one per added edge, instead of a heap of every cofacet. A
    // stream yields the cofacets of edge (i, j) in filtration order:
    //   phase A  diameter d(i, j): w with max(d(i,w), d(j,w)) <= d(i, j), by
    //            descending w (descending cofacet index);
    //   phase B  diameter max(d(i,w), d(j,w)) > d(i, j), ascending, from the
    //            rows of i and j sorted by distance; equal diameters by
    //            descending w; stops past the threshold.
    // Only the entries up to the current pivot are ever generated.
    struct EdgeStream {
        u32 i = 0, j = 0;
        f32 d = 0;
        u32 w = 0;                 // phase A: candidates below w remain
        u32 pi = 0, pj = 0;        // phase B cursors into sorted rows
        int phase = 0;             // 0 A, 1 B, 2 done
        bool has_pending = false;  // phase B one-entry lookahead
        u32 pend_w = 0;
        f32 pend_k = 0;
        std::vector<u32> tie;      // phase B group of equal diameters (descending w)
        u32 tie_pos = 0;
        f32 tie_k = 0;
        S head{};
    };

    // row v's other vertices by (distance, id), and those distances alongside
    // (read in order by the merge instead of looked up through the ids)
    const std::vector<u32>& sorted_row(u32 v) {
        if (srow_.empty()) {
            srow_.resize(n_);
            srowd_.resize(n_);
        }
        std::vector<u32>& r = srow_[v];
        if (r.empty() && n_ > 1) {
            r.reserve(n_ - 1);
            for (u32 w = 0; w < n_; ++w)
                if (w != v) r.push_back(w);
            const f32* row = cx_.mat + u64(v) * n_;
            std::sort(r.begin(), r.end(), [row](u32 a, u32 b) { return row[a] < row[b] || (row[a] == row[b] && a < b); });
            std::vector<f32>& rd = srowd_[v];
            rd.resize(r.size());
            for (size_t t = 0; t < r.size(); ++t) rd[t] = row[r[t]];
        }
        return r;
    }

    I tri_index(u32 i, u32 j, u32 w) const {   // i > j
        if (w > i) return B_(w, 3) + B_(i, 2) + B_(j, 1);
        if (w > j) return B_(i, 3) + B_(w, 2) + B_(j, 1);
        return B_(i, 3) + B_(j, 2) + B_(w, 1);
    }

    // next phase-B candidate (w, key) in ascending key, or false
    bool next_b(EdgeStream& s, u32& w_out, f32& k_out) {
        const std::vector<u32>& ri = sorted_row(s.i);
        const std::vector<u32>& rj = sorted_row(s.j);
        const f32* sdi = srowd_[s.i].data();
        const f32* sdj = srowd_[s.j].data();
        const f32* di = cx_.mat + u64(s.i) * n_;
        const f32* dj = cx_.mat + u64(s.j) * n_;
        for (;;) {
            const f32 ai = s.pi < ri.size() ? sdi[s.pi] : INFINITY;
            const f32 aj = s.pj < rj.size() ? sdj[s.pj] : INFINITY;
            if (s.pi >= ri.size() && s.pj >= rj.size()) return false;
            if (s.pi < ri.size() && (s.pj >= rj.size() || ai <= aj)) {
                const u32 w = ri[s.pi++];
                if (w == s.j) continue;
                if (dj[w] <= ai) {
                    w_out = w;
                    k_out = ai;
                    return true;
                }
            } else {
                const u32 w = rj[s.pj++];
                if (w == s.i) continue;
                if (di[w] < aj) {
                    w_out = w;
                    k_out = aj;
                    return true;
                }
            }
        }
    }

    // first position in a sorted row with distance > d
    u32 first_above(u32 v, f32 d) {
        sorted_row(v);
        const std::vector<f32>& rd = srowd_[v];
        return u32(std::upper_bound(rd.begin(), rd.end(), d) - rd.begin());
    }

    bool advance(EdgeStream& s) {
        const f32* di = cx_.mat + u64(s.i) * n_;
        const f32* dj = cx_.mat + u64(s.j) * n_;
        if (s.phase == 0) {
            while (s.w > 0) {
                const u32 w = --s.w;
                if (w == s.i || w == s.j) continue;
                const f32 a = di[w], b = dj[w];
                if ((a < b ? b : a) <= s.d) {
                    s.head = S{tri_index(s.i, s.j, w), s.d};
                    return true;
                }
            }
            s.phase = 1;
            s.pi = first_above(s.i, s.d);
            s.pj = first_above(s.j, s.d);
            s.has_pending = false;
            s.tie.clear();
            s.tie_pos = 0;
        }
        if (s.phase == 1) {
            if (s.tie_pos < s.tie.size()) {
                s.head = S{tri_index(s.i, s.j, s.tie[s.tie_pos++]), s.tie_k};
                return true;
            }
            u32 w;
            f32 k;
            if (s.has_pending) {
                w = s.pend_w;
                k = s.pend_k;
                s.has_pending = false;
            } else if (!next_b(s, w, k)) {
                s.phase = 2;
                return false;
            }
            if (!(k <= cx_.threshold)) {
                s.phase = 2;
                return false;
            }
            s.tie.clear();
            s.tie.push_back(w);
            u32 w2;
            f32 k2;
            while (next_b(s, w2, k2)) {
                if (k2 == k) {
                    s.tie.push_back(w2);
                } else {
                    s.has_pending = true;
                    s.pend_w = w2;
                    s.pend_k = k2;
                    break;
                }
            }
            if (s.tie.size() > 1) std::sort(s.tie.begin(), s.tie.end(), [](u32 a, u32 b) { return a > b; });
            s.tie_k = k;
            s.tie_pos = 1;
            s.head = S{tri_index(s.i, s.j, s.tie[0]), k};
            return true;
        }
        return false;
    }

    // the filtration order as one integer: diameter bits (non-negative
    // floats order as their bits), then the index descending
    static u128 order_key(const S& e) { return (u128(std::bit_cast<u32>(e.d)) << 64) | u128(~u64(e.idx)); }
    static S key_simplex(u128 k) { return S{I(~u64(k)), std::bit_cast<f32>(u32(k >> 64))}; }

    void lazy_push_stream(const S& edge) {
        u32 id;
        if (free_.empty()) {
            id = u32(pool_.size());
            pool_.emplace_back();
        } else {
            id = free_.back();
            free_.pop_back();
        }
        EdgeStream& s = pool_[id];
        decode_edge(u64(edge.idx), s.i, s.j);
        s.d = edge.d;
        s.w = n_;
        s.phase = 0;
        s.has_pending = false;
        s.tie.clear();
        s.tie_pos = 0;
        // An added column's coboundary starts at the current pivot, so the
        // entries of its edges' streams below the pivot cancel among
        // themselves: skip them (this also keeps the merge monotone).
        bool live = advance(s);
        while (live && order_key(s.head) < lheap_.last()) live = advance(s);
        if (live) {
            lheap_.push(order_key(s.head), id);
        } else {
            free_.push_back(id);
        }
    }

    void lazy_step(u32 id) {
        if (advance(pool_[id])) {
            lheap_.push(order_key(pool_[id].head), id);
        } else {
            free_.push_back(id);
        }
    }

    // earliest surviving entry after mod-2 cancellation (left in the merge)
    S lazy_pivot() {
        while (!lheap_.empty()) {
            const u128 key = lheap_.top_key();
            const u32 top = lheap_.pop();
            group_.clear();
            while (lheap_.has_last()) group_.push_back(lheap_.pop_last());   // same simplex
            if (group_.size() % 2 == 0) {   // top survives
                for (u32 g : group_) lazy_step(g);
                lheap_.push(key, top);      // equal to `last`: monotone
                return key_simplex(key);
            }
            lazy_step(top);
            for (u32 g : group_) lazy_step(g);
        }
        return S{kNone<I>, 0.0f};
    }

    // zero_pivot_cofacet of an edge, cached by edge index
    I edge_zero_cofacet(u64 eidx, f32 d) {
        if (zc_.key.empty()) zc_.reset(1024);
        const u64 hit = zc_.find(I(eidx));
        if (hit != ~0ull) return I(hit);
        u32 i, j;
        decode_edge(eidx, i, j);
        const f32* di = cx_.mat + u64(i) * n_;
        const f32* dj = cx_.mat + u64(j) * n_;
        I found = kNone<I>;
        for (u32 w = n_; w-- > 0;) {
            if (w == i || w == j) continue;
            const f32 a = di[w], b = dj[w];
            if ((a < b ? b : a) <= d) {
                found = tri_index(i, j, w);
                break;
            }
        }
        zc_.put(I(eidx), u64(found));
        return found;
    }

    // (f, t) apparent with triangle t the cofacet: f, else none
    S apparent_edge_of_triangle(const S& t) {
        u32 v[3];
        decode(t.idx, 3, v);
        const f32 d01 = cx_.dist(v[0], v[1]), d02 = cx_.dist(v[0], v[2]), d12 = cx_.dist(v[1], v[2]);
        // zero pivot facet: facets in ascending index: drop v0 (v1 v2), drop v1 (v0 v2), drop v2 (v0 v1)
        u32 a, b;
        if (d12 == t.d) { a = v[1]; b = v[2]; }
        else if (d02 == t.d) { a = v[0]; b = v[2]; }
        else if (d01 == t.d) { a = v[0]; b = v[1]; }
        else return S{kNone<I>, 0.0f};
        const u64 e = u64(a) * (a - 1) / 2 + b;
        return edge_zero_cofacet(e, t.d) == t.idx ? S{I(e), t.d} : S{kNone<I>, 0.0f};
    }

    bool reduce_dim1_lazy() {
        if constexpr (!std::is_same_v<Cx, DenseView> || !std::is_same_v<I, u64>) {
            return false;
        } else {
            if (!cx_.full) return false;
            pivots_.reset(cols_.size());
            red_off_.assign(1, 0);
            red_.clear();
            u64 emergent = 0;
            std::vector<S> added;
            for (u64 c = 0; c < cols_.size(); ++c) {
                const S col = cols_[c];
                lheap_.clear();
                free_.clear();
                for (u32 id = 0; id < pool_.size(); ++id) free_.push_back(id);
                added.clear();
                // emergent pair: the first cofacet of equal diameter (phase A head)
                lazy_push_stream(col);
                if (!lheap_.empty() && key_simplex(lheap_.top_key()).d == col.d) {
                    const S cf = key_simplex(lheap_.top_key());
                    if (!pivots_.has(cf.idx) && apparent_edge_of_triangle(cf).idx == kNone<I>) {
                        pivots_.put(cf.idx, c);
                        ++emergent;
                        red_off_.push_back(red_.size());
                        continue;
                    }
                }
                S piv = lazy_pivot();
                for (;;) {
                    if (piv.idx == kNone<I>) {
                        out_.add(1, col.d, INFINITY, u128(col.idx), u128(kNone<I>));
                        break;
                    }
                    const u64 jc = pivots_.find(piv.idx);
                    if (jc != ~0ull) {
                        added.push_back(cols_[jc]);
                        lazy_push_stream(cols_[jc]);
                        for (u64 t = red_off_[jc]; t < red_off_[jc + 1]; ++t) {
                            added.push_back(red_[t]);
                            lazy_push_stream(red_[t]);
                        }
                        piv = lazy_pivot();
                        continue;
                    }
                    const S fa = apparent_edge_of_triangle(piv);
                    if (fa.idx != kNone<I>) {
                        added.push_back(fa);
                        lazy_push_stream(fa);
                        piv = lazy_pivot();
                        continue;
                    }
                    if (piv.d > col.d) out_.add(1, col.d, piv.d, u128(col.idx), u128(piv.idx));
                    pivots_.put(piv.idx, c);
                    std::sort(added.begin(), added.end(), [](const S& a, const S& b) { return a.idx < b.idx; });
                    for (size_t x = 0; x < added.size();) {
                        if (x + 1 < added.size() && added[x].idx == added[x + 1].idx) x += 2;
                        else red_.push_back(added[x++]);
                    }
                    break;
                }
                red_off_.push_back(red_.size());
            }
            tel_.columns.push_back(cols_.size());
            tel_.emergent.push_back(emergent);
            return true;
        }
    }

    std::vector<std::vector<u32>> srow_;
    std::vector<std::vector<f32>> srowd_;
    std::vector<EdgeStream> pool_;
    std::vector<u32> free_, group_;
    RadixHeap lheap_;
    PivotTable<I> zc_;

    // What a finished column j keeps for the later columns that meet its
    // pivot: nothing (an emergent pair: its coboundary is its own, enumerated
    // again), its reduction (the simplices added to it: their coboundaries
    // enumerated again), or, on sparse complexes, where coboundaries are
    // short, its reduced coboundary itself (pushed as it is).
    enum : u8 { KEEP_NONE, KEEP_REDUCTION, KEEP_COBOUNDARY };
    std::vector<u8> keep_;

    void reduce(u32 dim) {
        const u32 nv = dim + 1;
        // cofacets have nv + 1 vertices: indices below C(n, nv + 1)
        radix_ = std::is_same_v<I, u64> || B_(n_, nv + 1) <= I(~u64(0));
        constexpr bool store_cob = std::is_same_v<Cx, SparseView>;
        pivots_.reset(cols_.size());
        red_off_.assign(1, 0);
        red_.clear();
        keep_.assign(cols_.size(), KEEP_NONE);
        u64 emergent = 0;
        std::vector<S> buf;
        std::vector<S> added;
        u32 vd[KM];
        for (u64 c = 0; c < cols_.size(); ++c) {
            const S col = cols_[c];
            heap_.clear();
            added.clear();
            buf.clear();
            decode(col.idx, nv, vd);
            S piv{kNone<I>, 0.0f};
            bool fresh = false, checking = true;
            cofacet_simplices(vd, nv, col.d, true, cx_.threshold, [&](const S& cf, u32, u32) {
                if (checking && cf.d == col.d) {
                    if (!pivots_.has(cf.idx) && apparent_facet_of(cf, nv + 1).idx == kNone<I>) {
                        piv = cf;           // emergent pair: the column is reduced already
                        fresh = true;
                        return false;
                    }
                    checking = false;
                }
                buf.push_back(cf);
                return true;
            });
            if (fresh) {
                ++emergent;
            } else {
                col_load(buf);
                piv = col_pivot();
            }
            for (;;) {
                if (piv.idx == kNone<I>) {
                    out_.add(dim, col.d, INFINITY, u128(col.idx), u128(kNone<I>));
                    break;
                }
                if (fresh) {
                    pivots_.put(piv.idx, c);
                    break;
                }
                const u64 j = pivots_.find(piv.idx);
                if (j != ~0ull) {
                    if (keep_[j] == KEEP_COBOUNDARY) {
                        for (u64 t = red_off_[j]; t < red_off_[j + 1]; ++t) col_push(red_[t]);
                    } else {
                        added.push_back(cols_[j]);
                        push_coboundary(cols_[j], nv);
                        for (u64 t = red_off_[j]; t < red_off_[j + 1]; ++t) {
                            added.push_back(red_[t]);
                            push_coboundary(red_[t], nv);
                        }
                    }
                    piv = col_pivot();
                    continue;
                }
                const S fa = apparent_facet_of(piv, nv + 1);
                if (fa.idx != kNone<I>) {
                    added.push_back(fa);
                    push_coboundary(fa, nv);
                    piv = col_pivot();
                    continue;
                }
                if (piv.d > col.d) out_.add(dim, col.d, piv.d, u128(col.idx), u128(piv.idx));
                pivots_.put(piv.idx, c);
                if constexpr (store_cob) {
                    // the reduced coboundary: the pivot and what follows it
                    keep_[c] = KEEP_COBOUNDARY;
                    col_drain([&](const S& e) { red_.push_back(e); });
                } else {
                    // the reduction column modulo 2
                    keep_[c] = KEEP_REDUCTION;
                    std::sort(added.begin(), added.end(), [](const S& a, const S& b) { return a.idx < b.idx; });
                    for (size_t x = 0; x < added.size();) {
                        if (x + 1 < added.size() && added[x].idx == added[x + 1].idx) x += 2;
                        else red_.push_back(added[x++]);
                    }
                }
                break;
            }
            red_off_.push_back(red_.size());
        }
        tel_.columns.push_back(cols_.size());
        tel_.emergent.push_back(emergent);
    }

    // columns of dimension dim + 1 from the simplices of dimension dim
    void assemble(u32 dim) {
        const u32 nv = dim + 1;
        const bool keep = dim + 1 < maxdim_;
        const u64 ns = simplices_.size();
        // Contiguous ranges of the simplices, one per thread; the outputs are
        // concatenated in range order, so columns and next simplices come out
        // in the serial order.
        const u32 nt = u32(std::max<u64>(1, std::min<u64>(threads_, ns / 4096)));
        std::vector<std::vector<S>> cols(nt), nexts(nt);
        std::vector<u64> skipped(nt, 0);
        auto range = [&](u32 t) {
            const u64 lo = ns * t / nt, hi = ns * (t + 1) / nt;
            std::vector<S>& oc = cols[t];
            std::vector<S>& on = nexts[t];
            u32 vd[KM], cv[KM + 1];
            for (u64 x = lo; x < hi; ++x) {
                const S& s = simplices_[x];
                decode(s.idx, nv, vd);
                cofacet_simplices(vd, nv, s.d, false, cx_.threshold, [&](const S& cf, u32 w, u32 t2) {
                    if (keep) on.push_back(cf);
                    if (pivots_.has(cf.idx)) return true;          // cleared
                    for (u32 i = 0; i < t2; ++i) cv[i] = vd[i];
                    cv[t2] = w;
                    for (u32 i = t2; i < nv; ++i) cv[i + 1] = vd[i];
                    if (apparent_cofacet_of(cf, cv, nv + 1).idx != kNone<I> ||
                        apparent_facet_of(cf, cv, nv + 1).idx != kNone<I>) {
                        ++skipped[t];
                        return true;
                    }
                    oc.push_back(cf);
                    return true;
                });
            }
        };
        if (nt == 1) {
            range(0);
        } else {
            std::vector<std::thread> pool;
            for (u32 t = 1; t < nt; ++t) pool.emplace_back(range, t);
            range(0);
            for (auto& th : pool) th.join();
        }
        cols_.clear();
        for (auto& v : cols) cols_.insert(cols_.end(), v.begin(), v.end());
        std::vector<S> next;
        for (auto& v : nexts) next.insert(next.end(), v.begin(), v.end());
        simplices_.swap(next);
        std::sort(cols_.begin(), cols_.end(), [](const S& a, const S& b) { return before(b, a); });
        u64 sk = 0;
        for (u64 v : skipped) sk += v;
        tel_.apparent.push_back(sk);
    }

    const Cx& cx_;
    const u32 n_, maxdim_;
    u32 threads_ = 1;
    const Binomials<I> B_;
    Diagram& out_;
    Telemetry& tel_;
    std::vector<S> cols_, simplices_, red_;
    std::vector<u64> red_off_;
    PivotTable<I> pivots_;
    Heap heap_;
};

}  // namespace tops::cob
