// SPDX-License-Identifier: MIT
// cohomology.hpp — the host half of the cohomology (coboundary) engine: the
// serial column reduction that remains once the GPU has done the
// per-simplex work.
//
// The filtration and the pairs are those of Ripser's persistent cohomology
// (U. Bauer, J. Appl. Comput. Topol. 2021), the algorithm behind the
// reference's coboundary engine (include/turbotops/rips.hpp:489-1032):
// simplices are named by the combinatorial number system and ordered by
// (diameter ascending, index descending) (rips.hpp:116-118); a column's
// coboundary is reduced mod 2 against earlier pivots; pairs that the
// filtration makes apparent are skipped; pivots of one dimension clear the
// columns of the next.
//
// What runs where:
//   GPU (csrc/edges.hip, sparse.hip)  dimension-0 edges in filtration order
//                                     with their apparent flags
//   host (dim0 below)                 union-find: H0 deaths, dimension-1
//                                     columns
//   GPU (csrc/columns.hip, through a ColumnSource)
//                                     each next dimension's columns (cofacets
//                                     neither cleared nor apparent, sorted),
//                                     and per column its emergent candidate
//   host (reduce below)               the column reduction itself; only
//                                     pivot-dependent decisions are made here
//
// Every finite pair also records its birth and death simplices (for the
// gradient of a diagram with respect to the points).
#pragma once

#include <algorithm>
#include <array>
#include <bit>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
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

// Pascal's triangle up to C(n, K): row k holds C(0..n, k). Construction fails
// with overflow_error when a value does not fit I.
template <class I>
class Binomials {
  public:
    Binomials(u32 n, u32 K) : rows_(K + 1, std::vector<I>(u64(n) + 1, I(0))) {
        for (u32 v = 0; v <= n; ++v) rows_[0][v] = 1;
        for (u32 k = 1; k <= K; ++k)
            for (u32 v = k; v <= n; ++v)
                if (__builtin_add_overflow(rows_[k][v - 1], rows_[k - 1][v - 1], &rows_[k][v]))
                    throw std::overflow_error("rips: simplex index overflow");
    }
    I operator()(u32 v, u32 k) const { return k < rows_.size() ? rows_[k][v] : I(0); }
    const I* row(u32 k) const { return rows_[k].data(); }

  private:
    std::vector<std::vector<I>> rows_;
};

// Simplex indices with up to maxdim + 2 vertices fit I, the all-ones value
// left free as the "none" sentinel.
template <class I>
bool index_fits(u32 n, u32 maxdim) {
    try {
        const Binomials<I> b(n, maxdim + 2);
        return b(n, maxdim + 2) != kNone<I>;
    } catch (const std::overflow_error&) {
        return false;
    }
}

// ---- complexes ----------------------------------------------------------------
// Both views answer:
//   dist(i, j)
//   coboundary(U, k, diam, cap, f): every cofacet U + {w} (w outside U, U: k
//     vertices descending) with cd = max(diam, d(u, w) for u in U) <= cap, in
//     ascending w, as f(index, cd) — indices from the table B
//   zpc_above(U, k, d, above): the largest w > above outside U with every
//     d(u, w) <= d (the vertex that makes U's zero-pivot cofacet), or ~0u
//
// The cofacet index of U + {w} with t vertices of U above w is
//   sum_{i<t} C(u_i, k+1-i) + C(w, k+1-t) + sum_{i>=t} C(u_i, k-i);
// as w ascends past u_{t-1} that vertex moves from the first sum (its term
// C(u, k+1-i)) to the last (C(u, k-i)), so both sums are kept running.
template <class I>
struct IndexWalk {
    const Binomials<I>& B;
    const u32* U;
    u32 k, t;
    I above_sum = 0, below_sum = 0;
    IndexWalk(const Binomials<I>& b, const u32* u, u32 kk) : B(b), U(u), k(kk), t(kk) {
        for (u32 i = 0; i < k; ++i) above_sum += B(U[i], k + 1 - i);
    }
    // index of U + {w}; w must not decrease between calls
    I at(u32 w) {
        while (t > 0 && U[t - 1] < w) {
            --t;
            above_sum -= B(U[t], k + 1 - t);
            below_sum += B(U[t], k - t);
        }
        return above_sum + B(w, k + 1 - t) + below_sum;
    }
};

constexpr u32 kNoVertex = ~0u;

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

    template <class I, class F>
    void coboundary(const Binomials<I>& B, const u32* U, u32 k, f32 diam, f32 cap, F&& f) const {
        if (!(diam <= cap)) return;
        IndexWalk<I> walk(B, U, k);
        if (!full) {
            for (u32 w = 0; w < n; ++w) {
                f32 cd = diam;
                bool ok = true;
                for (u32 i = 0; i < k && ok; ++i) {
                    if (U[i] == w) ok = false;
                    else cd = std::max(cd, dist(U[i], w));
                }
                if (ok && cd <= cap) f(walk.at(w), cd);
            }
            return;
        }
        // 64 candidates at a time: a bit mask of the w whose distance to every
        // vertex is within the cap, then the set bits in ascending order
        for (u32 base = 0; base < n; base += 64) {
            const u32 len = std::min<u32>(64, n - base);
            u64 pass = len == 64 ? ~0ull : ((u64(1) << len) - 1);
            for (u32 i = 0; i < k && pass; ++i) {
                const f32* r = mat + u64(U[i]) * n + base;
                u64 bad = 0;
                for (u32 j = 0; j < len; ++j) bad |= u64(!(r[j] <= cap)) << j;
                pass &= ~bad;
                if (U[i] >= base && U[i] < base + len) pass &= ~(u64(1) << (U[i] - base));
            }
            while (pass) {
                const u32 w = base + u32(std::countr_zero(pass));
                pass &= pass - 1;
                f32 cd = diam;
                for (u32 i = 0; i < k; ++i) cd = std::max(cd, mat[u64(U[i]) * n + w]);
                f(walk.at(w), cd);
            }
        }
    }

    u32 zpc_above(const u32* U, u32 k, f32 d, int64_t above) const {
        for (int64_t w = int64_t(n) - 1; w > above; --w) {
            bool ok = true;
            for (u32 i = 0; i < k && ok; ++i) ok = U[i] != u32(w) && dist(U[i], u32(w)) <= d;
            if (ok) return u32(w);
        }
        return kNoVertex;
    }
};

// A weighted graph in CSR form (rows sorted by neighbour id, both
// directions, no loops): a simplex is a clique, its cofacets the common
// neighbours of its vertices.
struct SparseView {
    u32 n = 0;
    f32 threshold = INFINITY;
    const u64* row_ptr = nullptr;
    const u32* col = nullptr;
    const f32* w = nullptr;

    // position of j in row i, or ~0
    u64 find(u32 i, u32 j) const {
        const u32* b = col + row_ptr[i];
        const u32* e = col + row_ptr[i + 1];
        const u32* it = std::lower_bound(b, e, j);
        return (it != e && *it == j) ? u64(it - col) : ~0ull;
    }
    f32 dist(u32 i, u32 j) const {
        if (i == j) return 0.0f;
        const u64 p = find(i, j);
        return p == ~0ull ? INFINITY : w[p];
    }
    u32 shortest_row(const u32* U, u32 k) const {
        u32 best = 0;
        for (u32 i = 1; i < k; ++i)
            if (row_ptr[U[i] + 1] - row_ptr[U[i]] < row_ptr[U[best] + 1] - row_ptr[U[best]]) best = i;
        return best;
    }

    template <class I, class F>
    void coboundary(const Binomials<I>& B, const u32* U, u32 k, f32 diam, f32 cap, F&& f) const {
        // the shortest row drives in ascending order; every other row keeps a
        // cursor that only moves forward
        constexpr u32 KC = 34;
        u64 cur[KC], end[KC];
        for (u32 i = 0; i < k; ++i) {
            cur[i] = row_ptr[U[i]];
            end[i] = row_ptr[U[i] + 1];
        }
        const u32 drv = shortest_row(U, k);
        IndexWalk<I> walk(B, U, k);
        for (u64 p = cur[drv]; p < end[drv]; ++p) {
            const u32 x = col[p];
            f32 cd = std::max(diam, w[p]);
            bool ok = cd <= cap;
            for (u32 i = 0; i < k && ok; ++i) {
                if (i == drv) continue;
                while (cur[i] < end[i] && col[cur[i]] < x) ++cur[i];
                ok = cur[i] < end[i] && col[cur[i]] == x;
                if (ok) {
                    cd = std::max(cd, w[cur[i]]);
                    ok = cd <= cap;
                }
            }
            // a vertex of U is never its own neighbour, so x outside U
            // already follows from x being in every row
            if (ok) f(walk.at(x), cd);
        }
    }

    u32 zpc_above(const u32* U, u32 k, f32 d, int64_t above) const {
        const u32 drv = shortest_row(U, k);
        const u64 lo = row_ptr[U[drv]];
        for (u64 p = row_ptr[U[drv] + 1]; p > lo; --p) {
            const u32 x = col[p - 1];
            if (int64_t(x) <= above) break;
            if (!(w[p - 1] <= d)) continue;
            bool ok = true;
            for (u32 i = 0; i < k && ok; ++i) {
                if (i == drv) continue;
                if (U[i] == x) ok = false;
                else {
                    const u64 q = find(U[i], x);
                    ok = q != ~0ull && w[q] <= d;
                }
            }
            if (ok) return x;
        }
        return kNoVertex;
    }
};

// Position (in U, descending) of the vertex whose removal gives the
// zero-pivot facet — the first facet in ascending index order keeping an
// edge of length d — or k if there is none: the first position outside the
// intersection of the end points of the edges of length d.
template <class Cx>
u32 zpf_position(const Cx& cx, const u32* U, u32 k, f32 d) {
    u64 common = ~0ull;
    bool any = false;
    for (u32 a = 0; a < k; ++a)
        for (u32 b = a + 1; b < k; ++b)
            if (cx.dist(U[a], U[b]) == d) {
                common &= (u64(1) << a) | (u64(1) << b);
                any = true;
            }
    if (!any) return k;
    const u32 j = u32(std::countr_zero(~common));
    return j < k ? j : k;
}

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
    template <class F>
    void each_key(F&& f) const {
        for (const I& k : key)
            if (k != kNone<I>) f(k);
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

// ---- a set under symmetric difference ---------------------------------------------
// The simplices added to a column, modulo 2: toggling a member removes it.
// Open addressing over slots that point into a dense item array, so the
// members are read out as one array; removal moves the last item into the
// hole, and deleted slots are tombstones until the next rebuild.
template <class S>
class ParitySet {
  public:
    void clear() {
        if (!items_.empty() || tombs_) std::fill(slot_.begin(), slot_.end(), kEmpty);
        items_.clear();
        where_.clear();
        tombs_ = 0;
    }
    void toggle(const S& s) {
        if (slot_.empty() || 2 * (items_.size() + tombs_ + 1) > slot_.size()) rebuild();
        const u64 mask = slot_.size() - 1;
        u64 h = mix(s.idx) & mask;
        u64 hole = ~0ull;
        for (;; h = (h + 1) & mask) {
            const u32 e = slot_[h];
            if (e == kEmpty) break;
            if (e == kTomb) {
                if (hole == ~0ull) hole = h;
                continue;
            }
            if (items_[e].idx == s.idx) {   // present: remove
                slot_[h] = kTomb;
                ++tombs_;
                const u32 last = u32(items_.size() - 1);
                if (e != last) {
                    items_[e] = items_[last];
                    where_[e] = where_[last];
                    slot_[where_[e]] = e;
                }
                items_.pop_back();
                where_.pop_back();
                return;
            }
        }
        if (hole != ~0ull) {
            h = hole;
            --tombs_;
        }
        slot_[h] = u32(items_.size());
        items_.push_back(s);
        where_.push_back(h);
    }
    const std::vector<S>& items() const { return items_; }

  private:
    static constexpr u32 kEmpty = ~0u, kTomb = ~0u - 1;
    std::vector<u32> slot_;
    std::vector<S> items_;
    std::vector<u64> where_;
    u64 tombs_ = 0;
    template <class I>
    static u64 mix(I k) {
        u64 x = u64(k) + u64(k >> 32 >> 32) * 0xD6E8FEB86659FD93ull;
        x ^= x >> 32;
        x *= 0xD6E8FEB86659FD93ull;
        x ^= x >> 32;
        return x;
    }
    void rebuild() {
        u64 cap = 64;
        while (cap < 4 * (items_.size() + 1)) cap <<= 1;
        slot_.assign(cap, kEmpty);
        tombs_ = 0;
        for (u32 e = 0; e < items_.size(); ++e) {
            u64 h = mix(items_[e].idx) & (cap - 1);
            while (slot_[h] != kEmpty) h = (h + 1) & (cap - 1);
            slot_[h] = e;
            where_[e] = h;
        }
    }
};

// ---- where the columns come from -------------------------------------------
// Simplex indices cross this interface as 64-bit words, one or two per
// index (low word first), as the C ABI (include/ttgpu.h) passes them.
struct ColumnSource {
    virtual ~ColumnSource() = default;
    // per column (nv vertices, filtration values d): the zero-pivot cofacet
    // (cof) and its state: 0 none, 1 a candidate emergent pair (the column
    // pairs with it unless it is already a pivot), 2 not a candidate (the
    // cofacet has a zero-apparent facet)
    virtual void emergent(u32 nv, const u64* idx, const f32* d, u64 count, u64* cof, u8* state) = 0;
    // the simplices the next assemble starts from
    virtual void seed(u32 nv, const u64* idx, const f32* d, u64 count) = 0;
    // the columns of the next dimension from the seeded simplices, the
    // given pivots cleared, in reduction order, with their emergent states;
    // keep_next: the cofacets become the seeded simplices
    virtual void assemble(const u64* pivots, u64 npivots, bool keep_next, std::vector<u64>& idx,
                          std::vector<f32>& d, std::vector<u64>& cof, std::vector<u8>& state, u64& skipped) = 0;
};

// ---- the reducer ---------------------------------------------------------------
template <class I, class Cx>
class Cohomology {
  public:
    static constexpr u32 kWords = sizeof(I) / 8;
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

    // Dimension 0 from the GPU's edges (filtration order) and apparent flags:
    // a spanning forest by union-find gives the H0 deaths; the other edges
    // that are not apparent are the dimension-1 columns, latest first.
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
        cur_.clear();
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
                else cur_.push_back(S{I(idx[e]), d[e]});
            }
        }
        for (u32 v = 0; v < n_; ++v)
            if (find(v) == v) out_.add(0, 0.0f, INFINITY, u128(v), u128(kNone<I>));
        std::reverse(cur_.begin(), cur_.end());
        edges_ = idx;
        edge_d_ = d;
        edge_count_ = m;
        tel_.edges = m;
        tel_.apparent.push_back(skipped);
    }

    // Dimensions 1..maxdim: columns and emergent states from `src`.
    void run(ColumnSource& src) {
        if (maxdim_ >= 2) {
            std::vector<u64> w(edge_count_ * kWords, 0);
            for (u64 e = 0; e < edge_count_; ++e) w[e * kWords] = edges_[e];
            src.seed(2, w.data(), edge_d_, edge_count_);
        }
        std::vector<u64> wi, wc, piv;
        std::vector<f32> wd;
        for (u32 dim = 1; dim <= maxdim_; ++dim) {
            if (dim == 1) {
                to_words(cur_, wi);
                wd.resize(cur_.size());
                for (size_t c = 0; c < cur_.size(); ++c) wd[c] = cur_[c].d;
                wc.assign(cur_.size() * kWords, 0);
                state_.assign(cur_.size(), 0);
                src.emergent(2, wi.data(), wd.data(), cur_.size(), wc.data(), state_.data());
            } else {
                u64 skipped = 0;
                src.assemble(piv.data(), piv.size() / kWords, dim < maxdim_, wi, wd, wc, state_, skipped);
                tel_.apparent.push_back(skipped);
                cur_.resize(wd.size());
                for (size_t c = 0; c < wd.size(); ++c) cur_[c] = S{word_index(wi.data(), c), wd[c]};
            }
            cof_.resize(cur_.size());
            for (size_t c = 0; c < cur_.size(); ++c) cof_[c] = word_index(wc.data(), c);
            if (!(dim == 1 && reduce_dim1_lazy())) reduce(dim);
            if (dim < maxdim_) {
                piv.clear();
                pivots_.each_key([&](I k) {
                    piv.push_back(u64(k));
                    if (kWords == 2) piv.push_back(u64(k >> 32 >> 32));
                });
            }
        }
    }

    // ---- index arithmetic -------------------------------------------------
    // vertices of simplex `idx` with k vertices, descending: for each
    // position the largest v below the previous vertex with C(v, pos) <= idx
    void decode(I idx, u32 k, u32* U) const {
        u32 top = n_;
        for (u32 pos = k; pos >= 1; --pos) {
            const I* row = B_.row(pos);
            const u32 v = u32(std::upper_bound(row + pos - 1, row + top, idx) - row) - 1;
            U[k - pos] = v;
            idx -= row[v];
            top = v;
        }
    }
    // edge (a, b), a > b, of index a(a-1)/2 + b
    static void decode_edge(u64 idx, u32& a, u32& b) {
        u64 v = u64((1.0 + std::sqrt(1.0 + 8.0 * double(idx))) * 0.5);
        while (v * (v - 1) / 2 > idx) --v;
        while ((v + 1) * v / 2 <= idx) ++v;
        a = u32(v);
        b = u32(idx - v * (v - 1) / 2);
    }
    I index_of(const u32* U, u32 k) const {
        I s = 0;
        for (u32 i = 0; i < k; ++i) s += B_(U[i], k - i);
        return s;
    }

  private:
    static constexpr u32 KM = 34;

    static I word_index(const u64* w, size_t t) {
        if constexpr (kWords == 2) return (I(w[2 * t + 1]) << 32 << 32) | I(w[2 * t]);
        else return I(w[t]);
    }
    static void to_words(const std::vector<S>& v, std::vector<u64>& out) {
        out.assign(v.size() * kWords, 0);
        for (size_t t = 0; t < v.size(); ++t) {
            out[t * kWords] = u64(v[t].idx);
            if constexpr (kWords == 2) out[t * 2 + 1] = u64(v[t].idx >> 32 >> 32);
        }
    }

    // (F, P) is a zero-apparent pair with the simplex P (k vertices) the
    // cofacet: F = zpf(P) removes U[j], and no vertex above U[j] extends F
    // within P's diameter. Returns F, or none.
    S apparent_facet_of(const S& p, u32 k) const {
        u32 U[KM], F[KM];
        decode(p.idx, k, U);
        const u32 j = zpf_position(cx_, U, k, p.d);
        if (j >= k) return S{kNone<I>, 0.0f};
        for (u32 i = 0, o = 0; i < k; ++i)
            if (i != j) F[o++] = U[i];
        if (cx_.zpc_above(F, k - 1, p.d, int64_t(U[j])) != kNoVertex) return S{kNone<I>, 0.0f};
        return S{index_of(F, k - 1), p.d};
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
    Heap heap_;

    // the filtration order as one integer: diameter bits (non-negative
    // floats order as their bits), then the index descending
    static u128 order_key(const S& e) { return (u128(std::bit_cast<u32>(e.d)) << 64) | u128(~u64(e.idx)); }
    static S key_simplex(u128 k) { return S{I(~u64(k)), std::bit_cast<f32>(u32(k >> 64))}; }

    void col_clear() {
        if (radix_) col_.clear();
        else heap_.clear();
    }
    void col_push(const S& c) {
        if (!radix_) {
            heap_.push(c);
            return;
        }
        const u128 k = order_key(c);
        if (k >= col_.last()) col_.push(k, 0);
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

    void push_coboundary(const S& s, u32 k) {
        u32 U[KM];
        decode(s.idx, k, U);
        cx_.coboundary(B_, U, k, s.d, cx_.threshold, [&](I idx, f32 cd) { col_push(S{idx, cd}); });
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

    // zero-pivot cofacet of an edge, cached by edge index
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

    void keep_reduction() {
        for (const S& s : parity_.items()) red_.push_back(s);
    }

    bool reduce_dim1_lazy() {
        if constexpr (!std::is_same_v<Cx, DenseView> || !std::is_same_v<I, u64>) {
            return false;
        } else {
            if (!cx_.full) return false;
            pivots_.reset(cur_.size());
            red_off_.assign(1, 0);
            red_.clear();
            u64 emergent = 0;
            for (u64 c = 0; c < cur_.size(); ++c) {
                const S col = cur_[c];
                if (state_[c] == 1 && !pivots_.has(cof_[c])) {   // emergent pair
                    pivots_.put(cof_[c], c);
                    ++emergent;
                    red_off_.push_back(red_.size());
                    continue;
                }
                lheap_.clear();
                free_.clear();
                for (u32 id = 0; id < pool_.size(); ++id) free_.push_back(id);
                parity_.clear();
                lazy_push_stream(col);
                S piv = lazy_pivot();
                for (;;) {
                    if (piv.idx == kNone<I>) {
                        out_.add(1, col.d, INFINITY, u128(col.idx), u128(kNone<I>));
                        break;
                    }
                    const u64 jc = pivots_.find(piv.idx);
                    if (jc != ~0ull) {
                        parity_.toggle(cur_[jc]);
                        lazy_push_stream(cur_[jc]);
                        for (u64 t = red_off_[jc]; t < red_off_[jc + 1]; ++t) {
                            parity_.toggle(red_[t]);
                            lazy_push_stream(red_[t]);
                        }
                        piv = lazy_pivot();
                        continue;
                    }
                    const S fa = apparent_edge_of_triangle(piv);
                    if (fa.idx != kNone<I>) {
                        parity_.toggle(fa);
                        lazy_push_stream(fa);
                        piv = lazy_pivot();
                        continue;
                    }
                    if (piv.d > col.d) out_.add(1, col.d, piv.d, u128(col.idx), u128(piv.idx));
                    pivots_.put(piv.idx, c);
                    keep_reduction();
                    break;
                }
                red_off_.push_back(red_.size());
            }
            tel_.columns.push_back(cur_.size());
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

    // Columns in reduction order. A column whose emergent candidate is not
    // yet a pivot pairs with it at once (its coboundary's first entry of
    // equal diameter, pivot-free and with no apparent facet, is its pivot).
    // Otherwise the working coboundary is reduced: a pivot that belongs to
    // an earlier column adds that column; a pivot with a zero-apparent facet
    // adds the facet; any other pivot closes the column.
    void reduce(u32 dim) {
        const u32 k = dim + 1;
        radix_ = std::is_same_v<I, u64> || B_(n_, k + 1) <= I(~u64(0));
        constexpr bool store_cob = std::is_same_v<Cx, SparseView>;
        const u64 nc = cur_.size();
        pivots_.reset(nc);
        red_off_.assign(1, 0);
        red_.clear();
        keep_.assign(nc, KEEP_NONE);
        u64 emergent = 0;
        for (u64 c = 0; c < nc; ++c) {
            const S col = cur_[c];
            if (state_[c] == 1 && !pivots_.has(cof_[c])) {
                pivots_.put(cof_[c], c);
                ++emergent;
                red_off_.push_back(red_.size());
                continue;
            }
            col_clear();
            parity_.clear();
            push_coboundary(col, k);
            for (S piv = col_pivot();;) {
                if (piv.idx == kNone<I>) {
                    out_.add(dim, col.d, INFINITY, u128(col.idx), u128(kNone<I>));
                    break;
                }
                const u64 j = pivots_.find(piv.idx);
                if (j != ~0ull) {
                    if (keep_[j] == KEEP_COBOUNDARY) {
                        for (u64 t = red_off_[j]; t < red_off_[j + 1]; ++t) col_push(red_[t]);
                    } else {
                        parity_.toggle(cur_[j]);
                        push_coboundary(cur_[j], k);
                        for (u64 t = red_off_[j]; t < red_off_[j + 1]; ++t) {
                            parity_.toggle(red_[t]);
                            push_coboundary(red_[t], k);
                        }
                    }
                    piv = col_pivot();
                    continue;
                }
                const S fa = apparent_facet_of(piv, k + 1);
                if (fa.idx != kNone<I>) {
                    parity_.toggle(fa);
                    push_coboundary(fa, k);
                    piv = col_pivot();
                    continue;
                }
                if (piv.d > col.d) out_.add(dim, col.d, piv.d, u128(col.idx), u128(piv.idx));
                pivots_.put(piv.idx, c);
                if constexpr (store_cob) {
                    keep_[c] = KEEP_COBOUNDARY;
                    col_drain([&](const S& e) { red_.push_back(e); });
                } else {
                    keep_[c] = KEEP_REDUCTION;
                    keep_reduction();
                }
                break;
            }
            red_off_.push_back(red_.size());
        }
        tel_.columns.push_back(nc);
        tel_.emergent.push_back(emergent);
    }

    const Cx& cx_;
    const u32 n_, maxdim_;
    const Binomials<I> B_;
    Diagram& out_;
    Telemetry& tel_;
    std::vector<S> cur_, red_;        // columns of the current dimension; stored reductions
    std::vector<I> cof_;              // per column: emergent candidate
    std::vector<u8> state_;           // per column: emergent state
    std::vector<u64> red_off_;
    PivotTable<I> pivots_;
    ParitySet<S> parity_;
    const u64* edges_ = nullptr;      // dimension-0 edges (the first seed)
    const f32* edge_d_ = nullptr;
    u64 edge_count_ = 0;
};

}  // namespace tops::cob
