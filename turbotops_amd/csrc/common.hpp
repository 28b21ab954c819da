// SPDX-License-Identifier: MIT
// common.hpp — shared host/device plumbing of libttgpu: status handling,
// HIP error checks, device buffers, and the exact scalar distance kernels
// every HIP kernel of the library evaluates.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>

#include "powf.hpp"
#include "ttgpu.h"

namespace ttg {

using u8 = uint8_t;
using u32 = uint32_t;
using u64 = uint64_t;
using i64 = int64_t;

// ---- status -------------------------------------------------------------
struct Error : std::exception {
    tt_status code;
    std::string msg;
    Error(tt_status c, std::string m) : code(c), msg(std::move(m)) {}
    const char* what() const noexcept override { return msg.c_str(); }
};

[[noreturn]] inline void fail(tt_status c, const std::string& m) { throw Error(c, m); }

void set_error(const std::string& m);

// Runs f, mapping every exception to a status and the thread-local message.
template <typename F>
tt_status guarded(F&& f) {
    try {
        f();
        return TT_OK;
    } catch (const Error& e) {
        set_error(e.msg);
        return e.code;
    } catch (const std::bad_alloc&) {
        set_error("out of host memory");
        return TT_EDEVICE;
    } catch (const std::exception& e) {
        set_error(e.what());
        return TT_EDEVICE;
    }
}

#define TTG_HIP(call)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            ::ttg::fail(TT_EDEVICE, std::string("hip: ") + #call + ": " + hipGetErrorString(e_)); \
    } while (0)

#define TTG_LAUNCH_CHECK() TTG_HIP(hipGetLastError())

// ---- device buffers (RAII, synchronous free) ---------------------------
template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    explicit DevBuf(size_t count) { alloc(count); }
    void alloc(size_t count) {
        release();
        n = count;
        if (count) TTG_HIP(hipMalloc(&p, count * sizeof(T)));
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    ~DevBuf() { release(); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
};

inline hipStream_t as_stream(void* s) { return static_cast<hipStream_t>(s); }

// ---- metric kinds -------------------------------------------------------
// Kernel-level metric variants: the reference's Lp splits into integral
// exponents <= 8 (square-and-multiply, metric.hpp:176-181, :187) and the
// powf form.
enum MetricVariant : int { MV_L1 = 0, MV_L2 = 1, MV_LINF = 2, MV_LP_INT = 3, MV_LP_POW = 4 };

inline MetricVariant variant_of(int kind, float p) {
    switch (kind) {
        case TT_METRIC_L1: return MV_L1;
        case TT_METRIC_L2: return MV_L2;
        case TT_METRIC_LINF: return MV_LINF;
        case TT_METRIC_LP: return (p == std::floor(p) && p <= 8.0f) ? MV_LP_INT : MV_LP_POW;
        default: fail(TT_EINVAL, "metric: unknown kind " + std::to_string(kind));
    }
}

void check_metric(int kind, float p);
void check_points(const float* pts, u64 count, const char* what);

// ---- exact scalar distance pieces (device) ------------------------------
// These reproduce the reference's scalar (TURBOTOPS_NO_SIMD) kernels
// operation by operation; the library is compiled with -ffp-contract=off so
// every fused multiply-add below is an explicit one.

// std::max(m, v) == (m < v) ? v : m
__device__ __forceinline__ float fmax_ref(float m, float v) { return (m < v) ? v : m; }

// fmax_ref for a running maximum m that starts at +0 and is never NaN, over
// terms v that are >= +0 or NaN (|a - b|): one v_max_f32. IEEE maxNum returns
// m when v is a (quiet) NaN, as (m < v) ? v : m does, and the two agree on
// every other pair (no -0 can occur); the compare-and-select form costs two
// instructions per coordinate.
__device__ __forceinline__ float fmax_acc(float m, float v) { return __builtin_fmaxf(m, v); }

// kern::ipow (metric.hpp:176-181)
__device__ __forceinline__ float ipow_ref(float x, u32 p) {
    float r = 1.0f;
    for (; p; p >>= 1, x *= x)
        if (p & 1) r *= x;
    return r;
}

// powf for the Lp terms and root: pow_f32 in powf.hpp (table-driven, double
// precision inside, within one ulp of glibc's powf).

}  // namespace ttg
