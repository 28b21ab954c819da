// SPDX-License-Identifier: MIT
// zerodiff.hpp — H0 and H1 of a full Rips filtration in the boundary
// direction: reduction to zero differentials with active enumeration.
//
// The method is C. Kapulkin and N. Kershaw's RedZeD ("reduction to zero
// differentials"), the engine the reference picks by default for H1 of dense
// full-Rips input (include/turbotops/turbotops.hpp:121-162, select_mode; the
// engine is include/turbotops/redzed.hpp:1001-1411). The reference picks it
// only for maxdim = 1, and this file covers exactly that case, written from
// the method: its inputs are what the GPU front end already produces (the
// full distance matrix and every edge up to the threshold in filtration
// order), so the host does no distance or sort work.
//
// Filtration. Edges in (diameter, combinatorial index) order; edge rank r is
// its 1-based position. A triangle enters right after its youngest edge, the
// last of its three edges in that order. Any order that refines diameter
// gives the same diagram.
//
// Classes. H1 is held in a basis of generators, each named by the rank of the
// edge that created it. A tree edge (it joins two components; union-find)
// has class 0. A non-tree edge x starts as its own generator. For each
// edge z with a nonzero class (an "active" edge) S[z] is the sorted set of
// generators in that class, and G[g] is the sorted set of active edges whose
// class contains generator g.
//
// Triangles. t = {a, b, w} entering with x = (a, b) has boundary class
// beta = S[x] + S[(a, w)] + S[(b, w)] (mod 2). If beta != 0 its youngest
// generator j dies at diam(t): every class that contains j gets beta added
// (S[z] ^= beta for z in G[j]), which removes j from the basis, and the bar
// (diam(j), diam(t)) is recorded when it is not empty.
//
// Active enumeration. Only triangles with an active face besides x can carry
// a class other than x's own; they are found from the active edges at x's
// endpoints. Every other triangle entering with x has boundary S[x], so the
// first of them kills x's class and the rest are then zero: one witness is
// enough, and it is found by a word-parallel AND of the endpoints' rows of
// the radius graph (all edges of diameter <= diam(x)).
//
// Ready test: is x the youngest face of {a, b, w} (a < b, D = diam(x), and
// d(w, a), d(w, b) <= D)? A face of diameter D that comes after x in
// combinatorial order would be. Face {w, b} comes after {a, b} iff w > a, and
// face {a, w} iff w > b. So: w < a: ready; a < w < b: ready iff d(w, b) < D;
// w > b: ready iff d(w, a) < D and d(w, b) < D.
//
// Output: bars with their birth and death simplices (edge and triangle
// combinatorial indices), as the cohomology engine records them, so the
// critical edges of every bar are read the same way.
#pragma once

#include <algorithm>
#include <bit>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <utility>
#include <vector>
#if defined(__AVX2__)
#include <immintrin.h>
#endif

#include "cohomology.hpp"

namespace tops::zd {

using cob::f32;
using cob::u128;
using cob::u32;
using cob::u64;

// u32 key (> 0) -> u32 value (> 0). Open addressing with linear probing over
// one array of (key << 32 | value) words, multiplicative hashing, load at most
// 1/2, and deletion by backward shift (no tombstones).
class RankMap {
public:
    RankMap() { rehash(16); }
    u32 get(u32 k) const {
        for (size_t i = home(k);; i = (i + 1) & mask_) {
            const u64 s = slots_[i];
            if (u32(s >> 32) == k) return u32(s);
            if (!s) return 0;
        }
    }
    void put(u32 k, u32 v) {
        if (2 * (used_ + 1) > slots_.size()) rehash(2 * slots_.size());
        size_t i = home(k);
        while (slots_[i] && u32(slots_[i] >> 32) != k) i = (i + 1) & mask_;
        if (!slots_[i]) ++used_;
        slots_[i] = (u64(k) << 32) | v;
    }
    void erase(u32 k) {
        size_t i = home(k);
        while (true) {
            if (!slots_[i]) return;
            if (u32(slots_[i] >> 32) == k) break;
            i = (i + 1) & mask_;
        }
        // shift later members of the probe run back over the hole
        size_t hole = i;
        for (size_t j = (i + 1) & mask_; slots_[j]; j = (j + 1) & mask_) {
            const size_t h = home(u32(slots_[j] >> 32));
            // j's home lies cyclically in (hole, j]: it stays
            const bool stays = hole <= j ? (hole < h && h <= j) : (hole < h || h <= j);
            if (!stays) {
                slots_[hole] = slots_[j];
                hole = j;
            }
        }
        slots_[hole] = 0;
        --used_;
    }
    template <class F>
    void for_each(F&& f) const {
        for (u64 s : slots_)
            if (s) f(u32(s >> 32), u32(s));
    }

private:
    std::vector<u64> slots_;
    size_t mask_ = 0, used_ = 0;
    int shift_ = 0;
    size_t home(u32 k) const { return size_t((u64(k) * 0x9E3779B97F4A7C15ull) >> shift_); }
    void rehash(size_t cap) {
        std::vector<u64> old;
        old.swap(slots_);
        slots_.assign(cap, 0);
        mask_ = cap - 1;
        shift_ = 64 - std::countr_zero(cap);
        used_ = 0;
        for (u64 s : old)
            if (s) put(u32(s >> 32), u32(s));
    }
};

// out = a ^ b for sorted sets
inline void xor_sorted(const std::vector<u32>& a, const std::vector<u32>& b, std::vector<u32>& out) {
    out.clear();
    size_t i = 0, j = 0;
    while (i < a.size() && j < b.size()) {
        if (a[i] < b[j]) out.push_back(a[i++]);
        else if (b[j] < a[i]) out.push_back(b[j++]);
        else {
            ++i;
            ++j;
        }
    }
    out.insert(out.end(), a.begin() + i, a.end());
    out.insert(out.end(), b.begin() + j, b.end());
}

// x ^= v (sorted); a single toggle when v holds one element
inline void xor_into(std::vector<u32>& x, const std::vector<u32>& v, std::vector<u32>& tmp) {
    if (v.size() == 1) {
        const auto it = std::lower_bound(x.begin(), x.end(), v[0]);
        if (it != x.end() && *it == v[0]) x.erase(it);
        else x.insert(it, v[0]);
        return;
    }
    if (v.size() * 16 <= x.size()) {
        // a few elements into a large set (a generator's set of class slots
        // holds thousands on 16-D clouds): find each by binary search and
        // copy the runs between them whole
        tmp.resize(x.size() + v.size());
        u32* out = tmp.data();
        const u32* base = x.data();
        size_t i = 0;
        for (u32 y : v) {
            const size_t p = size_t(std::lower_bound(x.begin() + i, x.end(), y) - x.begin());
            std::memcpy(out, base + i, (p - i) * sizeof(u32));
            out += p - i;
            i = p;
            if (i < x.size() && base[i] == y) ++i;   // present: it cancels
            else *out++ = y;                         // absent: it joins
        }
        std::memcpy(out, base + i, (x.size() - i) * sizeof(u32));
        out += x.size() - i;
        tmp.resize(size_t(out - tmp.data()));
        x.swap(tmp);
        return;
    }
    xor_sorted(x, v, tmp);
    x.swap(tmp);
}

inline u64 comb2(u32 lo, u32 hi) { return u64(hi) * (hi - 1) / 2 + lo; }

inline u128 comb3(u32 a, u32 b, u32 c) {
    u32 v[3] = {a, b, c};
    std::sort(v, v + 3);
    return u128(v[2]) * (v[2] - 1) * (v[2] - 2) / 6 + u128(v[1]) * (v[1] - 1) / 2 + v[0];
}

class BoundaryH1 {
public:
    // mat: full n x n distance matrix, row stride ld.
    BoundaryH1(u32 n, const f32* mat, u64 ld) : n_(n), W_((u64(n) + 63) / 64), mat_(mat), ld_(ld) {}

    // verts: 2 m vertex ids (lo < hi per edge), diam: m diameters, in
    // (diameter, combinatorial index) order, every edge up to the threshold.
    // witness_bit: bit 31 of each hi id is the edge's ready-witness flag (a
    // triangle with the edge as its youngest face exists; the GPU computes
    // it). With the flags an edge with no active face around it, by far the
    // most common case on most clouds, needs only its flag, and the radius
    // graph is not kept: adjacency at radius D is d(u, w) <= D read from the
    // matrix. Where edges with active faces around them become common (long
    // cycles: a noisy circle), the engine switches the radius graph on, so
    // their witness searches run word-parallel again.
    void run(u64 m, const u32* verts, const f32* diam, cob::Diagram& out, bool witness_bit = false) {
        if (m >= 0xffffffffull) throw std::length_error("boundary engine: more than 2^32 - 2 edges");
        out.ensure(1);
        m_ = m;
        verts_ = verts;
        diam_ = diam;
        flags_ = witness_bit;
        bits_ = !flags_;
        win_edges_ = win_cand_ = quiet_ = 0;
        adj_.clear();
        if (bits_) adj_.assign(n_ * W_, 0);
        seen_.assign(W_, 0);
        seed_a_.assign(n_, 0);
        seed_b_.assign(n_, 0);
        hseed_a_.assign(n_, 0);
        hseed_b_.assign(n_, 0);
        cand_.assign(n_, 0);
        stamp_.assign(n_, 0);
        ncand_ = 0;
        epoch_ = 0;
        inc_.assign(n_, {});
        parent_.resize(n_);
        for (u32 v = 0; v < n_; ++v) parent_[v] = v;
        size_.assign(n_, 1);
        G_ = RankMap();
        cls_.clear();
        head_.clear();
        owner_.clear();
        pos_lo_.clear();
        pos_hi_.clear();
        cls_free_.clear();
        gen_.clear();
        gen_free_.clear();
        gcnt_.clear();
        rptr_ = resolved_ = 0;
        linked_ = 0;
        for (u64 e = 0; e < m; ++e) {
            if (flags_ && !linked_) {
                // No edge is active, so a flagged edge has no candidates and
                // its class dies as it is born; nor is it a tree edge (the
                // other two edges of its witness triangle come before it).
                // Skip such edges reading only their flag bits.
                while (e < m && (verts[2 * e + 1] >> 31)) ++e;
                if (e == m) break;
            }
            const u32 a = verts[2 * e], b = hi_of(e);
            // the same for one flagged edge whose two ends have no active edge
            // (still counted in the radius graph's switching window)
            if (flags_ && (verts[2 * e + 1] >> 31) && inc_[a].empty() && inc_[b].empty()) {
                note_edge(diam[e]);
                continue;
            }
            const f32 d = diam[e];
            u32 ra = find(a), rb = find(b);
            if (ra != rb) {
                if (size_[ra] < size_[rb]) std::swap(ra, rb);
                parent_[rb] = ra;
                size_[ra] += size_[rb];
                if (d > 0.0f) out.add(0, 0.0f, d, u128(a), u128(comb2(a, b)));
            } else {
                cycle_edge(u32(e + 1), a, b, d, out);
            }
        }
        for (u32 v = 0; v < n_; ++v)
            if (find(v) == v) out.add(0, 0.0f, INFINITY, u128(v), u128(~u64(0)));
        std::vector<u32> alive;
        G_.for_each([&](u32 g, u32) { alive.push_back(g); });
        std::sort(alive.begin(), alive.end());
        for (u32 g : alive)
            out.add(1, diam_[g - 1], INFINITY, u128(comb2(verts_[2 * (g - 1)], hi_of(g - 1))), u128(~u64(0)));
    }

    u64 triangles_resolved() const { return resolved_; }

private:
    static constexpr u32 MULTI = ~0u;   // head_ of a class with two or more generators

    u32 n_;
    u64 W_;
    const f32* mat_;
    u64 ld_;
    u64 m_ = 0, rptr_ = 0, resolved_ = 0;
    u64 linked_ = 0;                     // active edges (classes linked into inc_)
    bool flags_ = false, bits_ = true;   // witness flags given; radius graph kept
    u32 win_edges_ = 0, win_cand_ = 0;   // cycle edges / witness searches, this window
    u32 quiet_ = 0;                      // quiet windows in a row with the radius graph on
    const u32* verts_ = nullptr;
    const f32* diam_ = nullptr;
    std::vector<u64> adj_;                               // radius graph rows, W_ words each
    std::vector<u64> seen_;                              // candidate vertices (for witness searches)
    // candidates of the current edge, t < ncand_: the third vertex, 1 + class
    // slot of edge (a, w) / (b, w) (0 = not active), their heads when collected
    std::vector<u32> cand_, seed_a_, seed_b_, hseed_a_, hseed_b_;
    u32 ncand_ = 0;
    // per vertex: (epoch << 32) | candidate index, for the current edge's epoch
    std::vector<u64> stamp_;
    u64 epoch_ = 0;
    // per vertex, its active edges: other end, class slot, and the class's
    // head word, kept current (set_head), so that collecting the candidates
    // reads the classes in order instead of one scattered load each
    struct Inc {
        u32 w, c, h;
    };
    std::vector<std::vector<Inc>> inc_;
    std::vector<u32> pos_lo_, pos_hi_;                   // per class slot: its entries in inc_[lo], inc_[hi]
    std::vector<u32> parent_, size_;
    // classes of active edges, by slot: generator set, its single element (0
    // when empty, MULTI when larger), and the edge's rank
    std::vector<std::vector<u32>> cls_;
    std::vector<u32> head_, owner_, cls_free_;
    // G sets (class slots holding a generator), by slot; G_: generator -> 1 + slot.
    // A G set is kept as a list that only grows between compactions: a slot
    // is appended when its class gains the generator and left in place when
    // the class loses it; the class's own sorted set decides membership, and
    // gcnt_ counts the true members. (Summing into sorted G sets copied
    // thousands of slots per resolve on 16-D clouds.)
    std::vector<std::vector<u32>> gen_;
    std::vector<u32> gen_free_, gcnt_, gys_;
    RankMap G_;
    std::vector<u32> beta_, touched_, tmp_;
    u32 cur_ = 0;                                        // the edge being processed (not linked yet)

    f32 dist(u32 u, u32 v) const { return mat_[u64(u) * ld_ + v]; }
    u32 hi_of(u64 e) const { return verts_[2 * e + 1] & 0x7fffffffu; }
    bool witness_of(u64 e) const { return verts_[2 * e + 1] >> 31; }
    u32 find(u32 x) {
        while (parent_[x] != x) {
            parent_[x] = parent_[parent_[x]];
            x = parent_[x];
        }
        return x;
    }
    // in the radius graph at radius D (all edges of diameter <= D)
    bool adjacent(u32 u, u32 w, f32 D) const {
        return bits_ ? (adj_[u * W_ + (w >> 6)] >> (w & 63)) & 1 : dist(u, w) <= D;
    }
    // Witness searches (edges whose class survived a seeded triangle) cost a
    // word-parallel AND with the radius graph, or two matrix rows without
    // it. Per window of 4096 cycle edges: many searches switch the radius
    // graph on (caught up to radius D), two quiet windows in a row switch it
    // off again (keeping it costs two bit writes per edge).
    void note_edge(f32 D) {
        if (++win_edges_ < 4096) return;
        if (!bits_ && win_cand_ > 64) {
            adj_.assign(n_ * W_, 0);
            rptr_ = 0;
            bits_ = true;
            advance(D);
            quiet_ = 0;
        } else if (bits_ && win_cand_ < 4 && ++quiet_ >= 2) {
            bits_ = false;
            adj_.clear();
            adj_.shrink_to_fit();
        } else if (win_cand_ >= 4) {
            quiet_ = 0;
        }
        win_edges_ = win_cand_ = 0;
    }
    void advance(f32 D) {
        if (!bits_) return;
        while (rptr_ < m_ && diam_[rptr_] <= D) {
            const u32 u = verts_[2 * rptr_], v = hi_of(rptr_);
            adj_[u * W_ + (v >> 6)] |= u64(1) << (v & 63);
            adj_[v * W_ + (u >> 6)] |= u64(1) << (u & 63);
            ++rptr_;
        }
    }
    bool ready(u32 a, u32 b, f32 D, u32 w) const {
        if (w < a) return true;
        if (w < b) return dist(w, b) < D;
        return dist(w, a) < D && dist(w, b) < D;
    }
    // a ready triangle {a, b, w} with no active face besides (a, b), or -1.
    // Any common neighbour w < a is ready without a distance read, so the
    // scan goes down from a's word first (nearest first: with labels that
    // follow the geometry, common neighbours sit near a), then up.
    int64_t leftover(u32 a, u32 b, f32 D) {
        ++win_cand_;
        if (!bits_) return leftover_dist(a, b, D);
        const u64* ra = &adj_[a * W_];
        const u64* rb = &adj_[b * W_];
        const u64 ta = a >> 6;
        const u64 below = (u64(1) << (a & 63)) - 1;   // bits of word ta under a
        for (u64 t = ta + 1; t-- > 0;) {
            u64 c = ra[t] & rb[t] & ~seen_[t];
            if (t == ta) c &= below;
            if (c) return int64_t(t * 64 + 63 - std::countl_zero(c));
        }
        for (u64 t = ta; t < W_; ++t) {
            u64 c = ra[t] & rb[t] & ~seen_[t];
            if (t == ta) c &= ~below & ~(u64(1) << (a & 63));
            while (c) {
                const u32 w = u32(t * 64 + std::countr_zero(c));
                c &= c - 1;
                if (ready(a, b, D, w)) return w;
            }
        }
        return -1;
    }

    // leftover() from the matrix rows: w below a first (ready once a common
    // neighbour), nearest first, then above
    int64_t leftover_dist(u32 a, u32 b, f32 D) const {
        const f32* ra = mat_ + u64(a) * ld_;
        const f32* rb = mat_ + u64(b) * ld_;
        auto unseen = [&](u32 w) { return !((seen_[w >> 6] >> (w & 63)) & 1); };
        for (u32 w = a; w-- > 0;)
            if (ra[w] <= D && rb[w] <= D && unseen(w)) return w;
        for (u32 w = a + 1; w < n_; ++w)
            if (w != b && ra[w] <= D && rb[w] <= D && unseen(w) && ready(a, b, D, w)) return w;
        return -1;
    }

    static u32 take(std::vector<std::vector<u32>>& pool, std::vector<u32>& freel) {
        if (freel.empty()) {
            pool.emplace_back();
            return u32(pool.size() - 1);
        }
        const u32 s = freel.back();
        freel.pop_back();
        return s;
    }
    void set_head(u32 c) {
        const auto& v = cls_[c];
        const u32 h = v.empty() ? 0 : v.size() == 1 ? v[0] : MULTI;
        head_[c] = h;
        const u32 z = owner_[c];
        if (h && z != cur_) {   // linked: refresh its two incident entries
            inc_[verts_[2 * (z - 1)]][pos_lo_[c]].h = h;
            inc_[hi_of(z - 1)][pos_hi_[c]].h = h;
        }
    }
    void link(u32 c, u32 lo, u32 hi) {
        ++linked_;
        pos_lo_[c] = u32(inc_[lo].size());
        inc_[lo].push_back({hi, c, head_[c]});
        pos_hi_[c] = u32(inc_[hi].size());
        inc_[hi].push_back({lo, c, head_[c]});
    }
    // class {x} and G[x] = {its slot} for a new non-tree edge x; returns the slot
    u32 create(u32 x) {
        const u32 c = take(cls_, cls_free_);
        if (head_.size() < cls_.size()) {
            head_.resize(cls_.size());
            owner_.resize(cls_.size());
        }
        if (pos_lo_.size() < cls_.size()) {
            pos_lo_.resize(cls_.size());
            pos_hi_.resize(cls_.size());
        }
        cls_[c].assign(1, x);
        head_[c] = x;
        owner_[c] = x;
        const u32 g = take(gen_, gen_free_);
        if (gcnt_.size() < gen_.size()) gcnt_.resize(gen_.size());
        gen_[g].assign(1, c);
        gcnt_[g] = 1;
        G_.put(x, g + 1);
        return c;
    }
    // removes class c's entry at position k of inc_[v]; the last entry moves
    // into its place, and its own position is updated
    void unlink_at(u32 v, u32 k) {
        auto& l = inc_[v];
        l[k] = l.back();
        l.pop_back();
        if (k < l.size()) {
            const Inc& mv = l[k];
            (v < mv.w ? pos_lo_ : pos_hi_)[mv.c] = k;
        }
    }
    void drop_class(u32 c) {
        const u32 z = owner_[c];
        cls_[c].clear();
        head_[c] = 0;
        cls_free_.push_back(c);
        if (z != cur_) {
            --linked_;
            unlink_at(verts_[2 * (z - 1)], pos_lo_[c]);
            unlink_at(hi_of(z - 1), pos_hi_[c]);
        }
    }
    // beta_ (nonzero boundary class of a triangle of diameter D): its
    // youngest generator dies
    bool holds(u32 c, u32 y) const { return std::binary_search(cls_[c].begin(), cls_[c].end(), y); }
    // the live members of generator y's list, sorted, without repeats
    void members(u32 g, u32 y, std::vector<u32>& out) const {
        out.clear();
        for (u32 c : gen_[g])
            if (holds(c, y)) out.push_back(c);
        std::sort(out.begin(), out.end());
        out.erase(std::unique(out.begin(), out.end()), out.end());
    }
    void resolve(f32 D, u128 tri, cob::Diagram& out) {
        ++resolved_;
        const u32 j = beta_.back();
        const u32 gj = G_.get(j) - 1;
        members(gj, j, touched_);
        const size_t others = beta_.size() - 1;
        gys_.resize(others);
        for (size_t t = 0; t < others; ++t) gys_[t] = G_.get(beta_[t]) - 1;
        for (u32 c : touched_) {
            xor_into(cls_[c], beta_, tmp_);
            // the class drops j and toggles each other generator of beta
            for (size_t t = 0; t < others; ++t) {
                const u32 gy = gys_[t];
                if (holds(c, beta_[t])) {
                    gen_[gy].push_back(c);
                    ++gcnt_[gy];
                } else {
                    --gcnt_[gy];
                }
            }
            set_head(c);
            if (cls_[c].empty()) drop_class(c);
        }
        for (size_t t = 0; t < others; ++t) {
            const u32 y = beta_[t], gy = gys_[t];
            if (!gcnt_[gy]) {
                gen_[gy].clear();
                gen_free_.push_back(gy);
                G_.erase(y);
            } else if (gen_[gy].size() > 2 * size_t(gcnt_[gy]) + 16) {
                members(gy, y, tmp_);
                gen_[gy].swap(tmp_);
            }
        }
        gen_[gj].clear();
        gen_free_.push_back(gj);
        G_.erase(j);
        const f32 birth = diam_[j - 1];
        if (birth != D) out.add(1, birth, D, u128(comb2(verts_[2 * (j - 1)], hi_of(j - 1))), tri);
    }
    // three class heads (0 = no class) of one generator each that cancel in
    // pairs: the boundary class is zero without looking at the sets
    static bool singles_cancel(u32 x, u32 y, u32 z) {
        if (x == MULTI || y == MULTI || z == MULTI) return false;
        if (!x) return y == z;
        if (!y) return x == z;
        if (!z) return x == y;
        return false;   // three nonzero singles never cancel
    }
    // beta_ = sum of the classes in slots p, q, r (1 + slot, 0 = none);
    // false when it is zero. Classes of one generator, the common case, are
    // summed from head_ alone.
    bool boundary_class(u32 p, u32 q, u32 r) {
        const u32 hp = p ? head_[p - 1] : 0, hq = q ? head_[q - 1] : 0, hr = r ? head_[r - 1] : 0;
        if (hp != MULTI && hq != MULTI && hr != MULTI) {
            u32 v[3] = {hp, hq, hr};
            if (v[0] > v[1]) std::swap(v[0], v[1]);
            if (v[1] > v[2]) std::swap(v[1], v[2]);
            if (v[0] > v[1]) std::swap(v[0], v[1]);
            beta_.clear();
            // zeros are absent classes; equal neighbours cancel in pairs
            for (int k = 0; k < 3; ++k) {
                if (!v[k]) continue;
                if (k + 1 < 3 && v[k + 1] == v[k]) {
                    ++k;
                    continue;
                }
                beta_.push_back(v[k]);
            }
            return !beta_.empty();
        }
        static const std::vector<u32> none;
        xor_sorted(p ? cls_[p - 1] : none, q ? cls_[q - 1] : none, tmp_);
        xor_sorted(tmp_, r ? cls_[r - 1] : none, beta_);
        return !beta_.empty();
    }

    // a ready triangle {a, b, w} without an active face other than (a, b):
    // leftover() with the current edge's candidates masked out
    int64_t unseeded_witness(u32 a, u32 b, f32 D) {
        for (u32 t = 0; t < ncand_; ++t) seen_[cand_[t] >> 6] |= u64(1) << (cand_[t] & 63);
        const int64_t w = leftover(a, b, D);
        // clear: word by word when the candidates are many, else each one's word
        if (u64(ncand_) * 4 >= W_) std::memset(seen_.data(), 0, W_ * sizeof(u64));
        else for (u32 t = 0; t < ncand_; ++t) seen_[cand_[t] >> 6] = 0;
        return w;
    }

    // The first candidate t' >= t whose recorded heads and x's head hx do not
    // cancel in pairs (ncand_ if none): almost every candidate cancels once
    // the edge's first resolve has run, so they are tested 8 at a time.
    u32 next_uncancelled(u32 t, u32 hx) const {
#if defined(__AVX2__)
        const __m256i X = _mm256_set1_epi32(int(hx)), Z = _mm256_setzero_si256(),
                      M = _mm256_set1_epi32(int(MULTI));
        const __m256i xz = _mm256_cmpeq_epi32(X, Z), xm = _mm256_cmpeq_epi32(X, M);
        for (; t + 8 <= ncand_; t += 8) {
            const __m256i ha = _mm256_loadu_si256(reinterpret_cast<const __m256i*>(&hseed_a_[t]));
            const __m256i hb = _mm256_loadu_si256(reinterpret_cast<const __m256i*>(&hseed_b_[t]));
            const __m256i pairs = _mm256_or_si256(
                _mm256_or_si256(_mm256_and_si256(xz, _mm256_cmpeq_epi32(ha, hb)),
                                _mm256_and_si256(_mm256_cmpeq_epi32(ha, Z), _mm256_cmpeq_epi32(X, hb))),
                _mm256_and_si256(_mm256_cmpeq_epi32(hb, Z), _mm256_cmpeq_epi32(X, ha)));
            const __m256i multi = _mm256_or_si256(xm, _mm256_or_si256(_mm256_cmpeq_epi32(ha, M),
                                                                      _mm256_cmpeq_epi32(hb, M)));
            const u32 keep = ~u32(_mm256_movemask_ps(_mm256_castsi256_ps(_mm256_andnot_si256(multi, pairs)))) & 0xffu;
            if (keep) return t + u32(std::countr_zero(keep));
        }
#endif
        for (; t < ncand_; ++t)
            if (!singles_cancel(hx, hseed_a_[t], hseed_b_[t])) return t;
        return ncand_;
    }

    void cycle_edge(u32 r, u32 a, u32 b, f32 D, cob::Diagram& out) {
        advance(D);
        // candidates in scan order (b's active edges, then a's), written
        // sequentially; a vertex met on both sides is found by its stamp
        const u64 ep = ++epoch_ << 32;
        u32 k = 0;
        for (const Inc& e : inc_[b])
            if (e.w != a && adjacent(a, e.w, D)) {
                stamp_[e.w] = ep | k;
                cand_[k] = e.w;
                seed_a_[k] = hseed_a_[k] = 0;
                seed_b_[k] = e.c + 1;
                hseed_b_[k] = e.h;
                ++k;
            }
        for (const Inc& e : inc_[a])
            if (e.w != b && adjacent(b, e.w, D)) {
                const u64 st = stamp_[e.w];
                u32 t = k;
                if ((st & ~u64(0xffffffffu)) == ep) {
                    t = u32(st);
                } else {
                    cand_[k] = e.w;
                    seed_b_[k] = hseed_b_[k] = 0;
                    ++k;
                }
                seed_a_[t] = e.c + 1;
                hseed_a_[t] = e.h;
            }
        ncand_ = k;
        cur_ = r;
        if (flags_) note_edge(D);
        const bool witness = flags_ ? witness_of(r - 1) : true;
        if (!ncand_) {
            // every triangle entering with x has boundary {x}: the first
            // ready one kills x's class as it is born
            if (flags_ ? !witness : leftover(a, b, D) < 0) {
                const u32 c = create(r);
                cur_ = 0;
                link(c, a, b);
            }
            cur_ = 0;
            return;
        }
        const u32 xc = create(r);
        bool dirty = false;
        // The heads recorded at collection stay current until a resolve
        // changes a class other than x's own (the usual resolve kills x's
        // fresh generator and touches x's class alone).
        bool recorded = true;
        for (u32 t = 0; t < ncand_; ++t) {
            if (recorded) {
                t = next_uncancelled(t, head_[xc]);
                if (t >= ncand_) break;
            }
            const u32 w = cand_[t];
            const u32 hx = head_[xc];
            const u32 ha = recorded ? hseed_a_[t] : seed_a_[t] ? head_[seed_a_[t] - 1] : 0;
            const u32 hb = recorded ? hseed_b_[t] : seed_b_[t] ? head_[seed_b_[t] - 1] : 0;
            if (singles_cancel(hx, ha, hb)) continue;
            if (!boundary_class(hx ? xc + 1 : 0, seed_a_[t], seed_b_[t])) continue;
            if (!ready(a, b, D, w)) continue;
            resolve(D, comb3(a, b, w), out);
            dirty = true;
            if (!(touched_.size() == 1 && touched_[0] == xc)) recorded = false;
        }
        if (head_[xc] && witness) {
            if (flags_ && !dirty) {
                // x's class is still {x}, so every seeded triangle had a
                // nonzero boundary and none was ready: the witness is an
                // unseeded triangle, and it kills x as it is born (a bar of
                // length 0, recorded by nobody)
                beta_ = cls_[xc];
                resolve(D, 0, out);
            } else {
                const int64_t w = unseeded_witness(a, b, D);
                if (w >= 0) {
                    beta_ = cls_[xc];
                    resolve(D, comb3(a, b, u32(w)), out);
                }
            }
        }
        cur_ = 0;
        if (head_[xc]) link(xc, a, b);
    }
};

}  // namespace tops::zd
