// SPDX-License-Identifier: MIT
// lib.hip — library-level entry points of libttgpu: errors, version,
// metric parsing and the dense layout decision.
#include <cstdlib>
#include <cstring>

#include "common.hpp"

namespace ttg {

namespace {
thread_local std::string g_last_error;
}

void set_error(const std::string& m) { g_last_error = m; }

// Metric::minkowski (metric.hpp:227-237): p must be >= 1 (NaN rejected).
void check_metric(int kind, float p) {
    if (kind < TT_METRIC_L1 || kind > TT_METRIC_LP)
        fail(TT_EINVAL, "metric: unknown kind " + std::to_string(kind));
    if (kind == TT_METRIC_LP && !(p >= 1.0f)) fail(TT_EINVAL, "metric: p must be in [1, inf]");
}

void check_points(const float* pts, u64 count, const char* what) {
    if (count && !pts) fail(TT_EINVAL, std::string(what) + ": null point array");
}

}  // namespace ttg

using namespace ttg;

extern "C" {

const char* tt_last_error(void) { return g_last_error.c_str(); }

const char* tt_version(void) { return "0.1.0"; }

int tt_device_count(void) {
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return c;
}

// Metric::parse (metric.hpp:240-253) with Metric::minkowski (:227-237).
tt_status tt_metric_parse(const char* metric, int* kind, float* p) {
    return guarded([&] {
        if (!metric || !kind || !p) fail(TT_EINVAL, "metric: null argument");
        const std::string s(metric);
        auto minkowski = [&](float pv) {
            if (!(pv >= 1.0f)) fail(TT_EINVAL, "metric: p must be in [1, inf]");
            *p = pv;
            if (pv == 1.0f) *kind = TT_METRIC_L1;
            else if (pv == 2.0f) *kind = TT_METRIC_L2;
            else if (std::isinf(pv)) *kind = TT_METRIC_LINF;
            else *kind = TT_METRIC_LP;
        };
        if (s == "l2" || s == "euclidean") return minkowski(2.0f);
        if (s == "l1" || s == "manhattan") return minkowski(1.0f);
        if (s == "linf" || s == "chebyshev") return minkowski(INFINITY);
        if (s.rfind("lp:", 0) == 0) {
            const std::string v = s.substr(3);
            if (v == "inf") return minkowski(INFINITY);
            char* end = nullptr;
            const float pv = std::strtof(v.c_str(), &end);
            if (end && *end == '\0' && !v.empty()) return minkowski(pv);
        }
        fail(TT_EINVAL, "metric: unknown '" + s + "' (expected l1 | l2 | linf | lp:<p>, p >= 1)");
    });
}

// dense_bytes and the make_dense layout gate (rips.hpp:384-386, :395-414).
tt_status tt_dense_layout(uint32_t n, uint64_t budget, int* layout) {
    return guarded([&] {
        if (!layout) fail(TT_EINVAL, "tt_dense_layout: null layout");
        const u64 full = u64(n) * n * sizeof(float);
        const u64 lower = u64(n) * (n > 0 ? n - 1 : 0) / 2 * sizeof(float);
        if (full <= budget) *layout = TT_LAYOUT_FULL;
        else if (lower <= budget) *layout = TT_LAYOUT_LOWER;
        else
            fail(TT_ELENGTH,
                 "rips: the dense distance matrix exceeds the memory budget; use a "
                 "finite threshold or a sparse graph");
    });
}

}  // extern "C"
