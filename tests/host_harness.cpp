// SPDX-License-Identifier: MIT
// host_harness.cpp — TEST INFRASTRUCTURE ONLY: a C ABI around the host
// reducers (turbotops_amd/host/reducer.hpp, the cohomology engine, and
// turbotops_amd/host/zerodiff.hpp, the boundary engine for H1) so the CPU
// test suite can check them
// against the reference with inputs the oracle computes (full matrix, sorted
// edges, apparent flags) — no GPU involved. Not part of the product.
#include <cstdlib>
#include <cstring>
#include <string>

#include "reducer.hpp"
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

template <class Cx>
int run(const Cx& cx, uint32_t maxdim, const uint64_t* idx, const float* d, const uint8_t* ap, uint64_t m,
        uint64_t* nbars, uint32_t** dims, float** b, float** e) {
    try {
        Diagram out;
        out.ensure(maxdim);
        Telemetry tel;
        // several threads, so the CPU tests also cover the parallel assembly
        if (index_fits<u64>(cx.n, maxdim)) {
            Cohomology<u64, Cx> r(cx, maxdim, out, tel);
            r.set_threads(4);
            r.dim0(idx, d, ap, m);
            r.run_from_dim1();
        } else {
            Cohomology<u128, Cx> r(cx, maxdim, out, tel);
            r.set_threads(4);
            r.dim0(idx, d, ap, m);
            r.run_from_dim1();
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
             const uint8_t* ap, uint64_t m, uint64_t* nbars, uint32_t** dims, float** b, float** e) {
    DenseView cx;
    cx.n = n;
    cx.threshold = threshold;
    cx.full = true;
    cx.mat = mat;
    return run(cx, maxdim, idx, d, ap, m, nbars, dims, b, e);
}

int hh_sparse(uint32_t n, const uint64_t* rp, const uint32_t* col, const float* w, float threshold, uint32_t maxdim,
              const uint64_t* idx, const float* d, const uint8_t* ap, uint64_t m, uint64_t* nbars, uint32_t** dims,
              float** b, float** e) {
    SparseView cx;
    cx.n = n;
    cx.threshold = threshold;
    cx.row_ptr = rp;
    cx.col = col;
    cx.w = w;
    return run(cx, maxdim, idx, d, ap, m, nbars, dims, b, e);
}

// apparent flags of sorted edges on a sparse complex (the host dim-0 test)
int hh_sparse_apparent(uint32_t n, const uint64_t* rp, const uint32_t* col, const float* w, uint32_t maxdim,
                       const uint64_t* idx, const float* d, uint64_t m, uint8_t* flags) {
    try {
        SparseView cx;
        cx.n = n;
        cx.row_ptr = rp;
        cx.col = col;
        cx.w = w;
        Diagram out;
        Telemetry tel;
        Cohomology<u64, SparseView> r(cx, maxdim, out, tel);
        for (uint64_t t = 0; t < m; ++t) flags[t] = r.edge_apparent(idx[t], d[t]);
        return 0;
    } catch (const std::exception& ex) {
        g_err = ex.what();
        return 1;
    }
}
}
