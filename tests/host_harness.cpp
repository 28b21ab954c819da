// SPDX-License-Identifier: MIT
// host_harness.cpp — TEST INFRASTRUCTURE ONLY: a C ABI around the host
// engines (turbotops_amd/host/cohomology.hpp, the cohomology engine's column
// reduction, and turbotops_amd/host/zerodiff.hpp, the boundary engine for H1)
// so the CPU test suite can check them against the reference with inputs the
// oracle computes (full matrix, sorted edges, apparent flags) — no GPU
// involved. The cohomology engine's columns and emergent states, which the
// product takes from the GPU (csrc/columns.hip), come here from the oracle's
// restatement (oracle/ttoracle.c, linked in). Not part of the product.
#include <cstdlib>
#include <cstring>
#include <string>

#include "cohomology.hpp"
#include "powf.hpp"
#include "ttoracle.h"
#include "zerodiff.hpp"

using namespace tops::cob;

namespace {
thread_local std::string g_err;

void flatten(const Diagram& dg, uint64_t* nbars, uint32_t** dims, float** b, float** e) {
    uint64_t k = 0;
    for (const auto& v : dg.birth) k += v.size();
    *nbars = k;
    *dims = static_cast<uint32_t*>(std::malloc(std::max<uint64_t>(1, k) * 4));
    *b = static_cast<float*>(std::malloc(std::max<uint64_t>(1, k) * 4));
    *e = static_cast<float*>(std::malloc(std::max<uint64_t>(1, k) * 4));
    uint64_t o = 0;
    for (uint32_t d = 0; d < dg.birth.size(); ++d)
        for (size_t i = 0; i < dg.birth[d].size(); ++i, ++o) {
            (*dims)[o] = d;
            (*b)[o] = dg.birth[d][i];
            (*e)[o] = dg.death[d][i];
        }
}

// The oracle's restatements of the GPU column stages, as a ColumnSource.
class OracleColumns final : public ColumnSource {
  public:
    OracleColumns(const DenseView& cx, u32 words) : sparse_(0), n_(cx.n), mat_(cx.mat), thr_(cx.threshold),
                                                    words_(words) {}
    OracleColumns(const SparseView& cx, u32 words)
        : sparse_(1), n_(cx.n), rp_(cx.row_ptr), col_(cx.col), w_(cx.w), thr_(cx.threshold), words_(words) {}
    void emergent(u32 nv, const u64* idx, const f32* d, u64 count, u64* cof, u8* state) override {
        tto_emergent_cx(sparse_, n_, mat_, rp_, col_, w_, thr_, nv, words_, idx, d, count, cof, state);
    }
    void seed(u32 nv, const u64* idx, const f32* d, u64 count) override {
        nv_ = nv;
        sidx_.assign(idx, idx + count * words_);
        sd_.assign(d, d + count);
    }
    void assemble(const u64* pivots, u64 npivots, bool keep_next, std::vector<u64>& idx, std::vector<f32>& d,
                  std::vector<u64>& cof, std::vector<u8>& state, u64& skipped) override {
        std::vector<u64> piv(pivots, pivots + npivots * words_);
        // the oracle takes the pivots sorted
        std::vector<u128> keys(npivots);
        for (u64 t = 0; t < npivots; ++t)
            keys[t] = words_ == 2 ? (u128(piv[2 * t + 1]) << 64) | piv[2 * t] : u128(piv[t]);
        std::sort(keys.begin(), keys.end());
        for (u64 t = 0; t < npivots; ++t) {
            if (words_ == 2) {
                piv[2 * t] = u64(keys[t]);
                piv[2 * t + 1] = u64(keys[t] >> 64);
            } else {
                piv[t] = u64(keys[t]);
            }
        }
        u64 nc = 0, nn = 0;
        u64 *ci = nullptr, *cc = nullptr, *ni = nullptr;
        f32 *cd = nullptr, *nd = nullptr;
        u8* cs = nullptr;
        tto_assemble_cx(sparse_, n_, mat_, rp_, col_, w_, thr_, nv_, words_, sidx_.data(), sd_.data(), sd_.size(),
                        piv.empty() ? nullptr : piv.data(), npivots, keep_next, &nc, &ci, &cd, &cc, &cs, &nn, &ni,
                        &nd, &skipped);
        idx.assign(ci, ci + nc * words_);
        cof.assign(cc, cc + nc * words_);
        d.assign(cd, cd + nc);
        state.assign(cs, cs + nc);
        if (keep_next) {
            sidx_.assign(ni, ni + nn * words_);
            sd_.assign(nd, nd + nn);
            ++nv_;
        }
        for (void* p : {(void*)ci, (void*)cc, (void*)ni, (void*)cd, (void*)nd, (void*)cs}) tto_free(p);
    }

  private:
    int sparse_;
    u32 n_;
    const f32* mat_ = nullptr;
    const u64* rp_ = nullptr;
    const u32* col_ = nullptr;
    const f32* w_ = nullptr;
    f32 thr_;
    u32 words_;
    u32 nv_ = 0;
    std::vector<u64> sidx_;
    std::vector<f32> sd_;
};

template <class Cx>
int run(const Cx& cx, uint32_t maxdim, const uint64_t* idx, const float* d, const uint8_t* ap, uint64_t m,
        int force_wide, uint64_t* nbars, uint32_t** dims, float** b, float** e) {
    try {
        Diagram out;
        out.ensure(maxdim);
        Telemetry tel;
        if (!force_wide && index_fits<u64>(cx.n, maxdim)) {
            Cohomology<u64, Cx> r(cx, maxdim, out, tel);
            r.dim0(idx, d, ap, m);
            OracleColumns src(cx, 1);
            r.run(src);
        } else {
            Cohomology<u128, Cx> r(cx, maxdim, out, tel);
            r.dim0(idx, d, ap, m);
            OracleColumns src(cx, 2);
            r.run(src);
        }
        flatten(out, nbars, dims, b, e);
        return 0;
    } catch (const std::exception& ex) {
        g_err = ex.what();
        return 1;
    }
}
}  // namespace

extern "C" {
const char* hh_last_error() { return g_err.c_str(); }

// H0 and H1 by the boundary engine: full n x n matrix and m edges as 2 m
// vertex ids (lo < hi) with diameters, in (diameter, comb index) order.
// witness_bit: bit 31 of each hi id carries the ready-witness flag.
int hh_boundary(const float* mat, uint32_t n, const uint32_t* verts, const float* d, uint64_t m, int witness_bit,
                uint64_t* nbars, uint32_t** dims, float** b, float** e) {
    try {
        Diagram out;
        out.ensure(1);
        tops::zd::BoundaryH1 eng(n, mat, n);
        if (n) eng.run(m, verts, d, out, witness_bit != 0);
        flatten(out, nbars, dims, b, e);
        return 0;
    } catch (const std::exception& ex) {
        g_err = ex.what();
        return 1;
    }
}
void hh_free(void* p) { std::free(p); }

int hh_dense(const float* mat, uint32_t n, float threshold, uint32_t maxdim, const uint64_t* idx, const float* d,
             const uint8_t* ap, uint64_t m, int force_wide, uint64_t* nbars, uint32_t** dims, float** b, float** e) {
    DenseView cx;
    cx.n = n;
    cx.threshold = threshold;
    cx.full = true;
    cx.mat = mat;
    return run(cx, maxdim, idx, d, ap, m, force_wide, nbars, dims, b, e);
}

int hh_sparse(uint32_t n, const uint64_t* rp, const uint32_t* col, const float* w, float threshold, uint32_t maxdim,
              const uint64_t* idx, const float* d, const uint8_t* ap, uint64_t m, int force_wide, uint64_t* nbars,
              uint32_t** dims, float** b, float** e) {
    SparseView cx;
    cx.n = n;
    cx.threshold = threshold;
    cx.row_ptr = rp;
    cx.col = col;
    cx.w = w;
    return run(cx, maxdim, idx, d, ap, m, force_wide, nbars, dims, b, e);
}


// The Lp kernels' power (turbotops_amd/csrc/powf.hpp), host build of the same
// source: out[t] = x[t]^p.
void hh_pow_f32(const float* x, uint64_t count, float p, float* out) {
    for (uint64_t t = 0; t < count; ++t) out[t] = ttg::pow_f32(x[t], p);
}
void hh_pow_half_f32(const float* x, uint64_t count, uint32_t n, float* out) {
    for (uint64_t t = 0; t < count; ++t) out[t] = ttg::pow_half_f32(x[t], n);
}

// The boundary engine's set and map primitives (turbotops_amd/host/
// zerodiff.hpp), for direct checks against plain references.
// xor_into: out = x ^ v for sorted sets; returns |out|.
uint64_t hh_xor_into(const uint32_t* x, uint64_t nx, const uint32_t* v, uint64_t nv, uint32_t* out) {
    std::vector<uint32_t> a(x, x + nx), b(v, v + nv), tmp;
    tops::zd::xor_into(a, b, tmp);
    std::copy(a.begin(), a.end(), out);
    return a.size();
}
// RankMap: ops[3t] = 0 put(key, val), 1 get(key) -> res[t], 2 erase(key).
void hh_rankmap(const uint32_t* ops, uint64_t count, uint32_t* res) {
    tops::zd::RankMap m;
    for (uint64_t t = 0; t < count; ++t) {
        const uint32_t op = ops[3 * t], k = ops[3 * t + 1], v = ops[3 * t + 2];
        res[t] = 0;
        if (op == 0) m.put(k, v);
        else if (op == 1) res[t] = m.get(k);
        else m.erase(k);
    }
}
}
